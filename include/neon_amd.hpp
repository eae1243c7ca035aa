// Neon-shaped C++ host API over the C ABI of neon_amd.h: Backend, dGrid /
// bGrid, Field, Container, Skeleton and the D3Q19 LBM containers, with the
// names, argument meanings and error behaviour of the reference's
// lid-driven-cavity benchmark path:
//   Neon::Backend                  libNeonSet/include/Neon/set/Backend.h
//   Neon::dGrid / Neon::bGrid      libNeonDomain/include/Neon/domain/{dGrid,bGrid}.h
//   Field::forEachActiveCell, updateDeviceData, updateHostData
//                                  libNeonDomain/include/Neon/domain/interface/FieldBase.h
//   Neon::set::Container::run      libNeonSet/include/Neon/set/Containter.h:25-35
//   Neon::skeleton::Skeleton       libNeonSkeleton/include/Neon/skeleton/Skeleton.h:56-65
//   LbmContainers, LbmIterationD3Q19
//                                  benchmarks/lbm-lid-driven-cavity-flow/src/{LbmTools.h,LbmIteration.h}
// Compiled with hipcc, the header also offers Neon's generic Container API on
// the dense grid (grid.newContainer, Loader::load, partition getNghData) for
// a caller's own stencils (end of this file).
// Header-only, C++17; link with libneon_amd.so.  One process drives one GPU:
// a Backend takes exactly one device id (the reference's multi-device
// partitioning collapses to one partition, so there is no halo exchange and
// the Occ / TransferMode / StencilSemantic options are accepted and have
// nothing to overlap or exchange).  Errors throw neon_amd::NeonException
// carrying nx_last_error(), where the reference throws Neon::NeonException.
#pragma once

#include "neon_amd.h"

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <functional>
#include <ostream>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

namespace neon_amd {

class NeonException : public std::runtime_error
{
   public:
    using std::runtime_error::runtime_error;
};

inline void check(int code, const std::string& where)
{
    if (code != NX_OK) throw NeonException(where + ": " + nx_last_error());
}

// index_3d is usable in device lambdas of hipcc-compiled callers
#if defined(__HIPCC__)
#define NX_HOST_DEVICE __host__ __device__
#else
#define NX_HOST_DEVICE
#endif

struct index_3d
{
    int32_t x = 0, y = 0, z = 0;
    NX_HOST_DEVICE index_3d() = default;
    NX_HOST_DEVICE index_3d(int32_t x_, int32_t y_, int32_t z_) : x(x_), y(y_), z(z_) {}
    NX_HOST_DEVICE size_t rMulTyped() const { return size_t(x) * size_t(y) * size_t(z); }
};

enum class Runtime { stream };
enum class Occ { none, standard };
enum class TransferMode { get, put };
enum class StencilSemantic { standard, streaming };
enum class IoFileType { ASCII, BINARY };

namespace vtk {

// Legacy VTK writer for one field over the full grid box, in the format of
// the reference's ioToVTK (libNeonCore/include/Neon/core/tools/io/ioToVTK.h:
// header :430-442, grid info :352-354, SCALARS / LOOKUP_TABLE / METADATA
// blocks :239-272, CELL_DATA before POINT_DATA :320-325; text values
// separated by one space with a newline per x row :194-215; binary values
// big-endian followed by one newline :166-183).  `values` holds
// box.x * box.y * box.z * card values, x fastest, then the component.
// isNodeSpace false (the reference's default for Field::ioToVtk) writes the
// values as cell data of a (box + 1)-point lattice.
template <typename V>
inline const char* typeName()
{
    if constexpr (std::is_same_v<V, double>) return "double";
    else if constexpr (std::is_same_v<V, float>) return "float";
    else if constexpr (std::is_same_v<V, int32_t> || std::is_same_v<V, uint32_t>) return "int";
    else static_assert(sizeof(V) == 0, "ioToVtk exports double, float or 32-bit integer values");
}

template <typename V>
inline void writeScalars(std::ostream& out, const std::string& name, const std::vector<V>& values,
                         const index_3d& box, int card, IoFileType type)
{
    out << "SCALARS " << name << " " << typeName<V>() << " ";
    if (card != 1) out << " " << card << "\n";
    else out << "\n";
    out << "LOOKUP_TABLE default\n";
    const size_t row = size_t(box.x) * size_t(card);
    const size_t rows = size_t(box.y) * size_t(box.z);
    if (type == IoFileType::ASCII) {
        for (size_t r = 0; r < rows; r++) {
            for (size_t i = 0; i < row; i++) out << values[r * row + i] << " ";
            out << "\n";
        }
    } else {
        std::vector<unsigned char> buf(row * sizeof(V));
        for (size_t r = 0; r < rows; r++) {
            for (size_t i = 0; i < row; i++) {
                unsigned char b[sizeof(V)];
                std::memcpy(b, &values[r * row + i], sizeof(V));
                for (size_t k = 0; k < sizeof(V); k++) buf[i * sizeof(V) + k] = b[sizeof(V) - 1 - k];
            }
            out.write(reinterpret_cast<const char*>(buf.data()), std::streamsize(buf.size()));
        }
        out << "\n";
    }
    out << "METADATA\n";
    out << "INFORMATION 0\n\n";
}

// The whole file: fileName + ".vtk" holding one field of `card` components
// over the box `dim` (values x fastest, then the component), plus the
// all-zero "Domain" cell field when includeDomain (ioToVTK.h:413-445,
// IOGridVTK_imp.h:6-18, FieldBase_imp.h:297-318).
template <typename V>
inline void writeFile(const std::string& fileName, const std::string& fieldName, const std::vector<V>& values,
                      const index_3d& dim, int card, bool includeDomain, IoFileType type, bool isNodeSpace)
{
    const index_3d node = isNodeSpace ? dim : index_3d(dim.x + 1, dim.y + 1, dim.z + 1);
    const index_3d vox(node.x - 1, node.y - 1, node.z - 1);
    std::ofstream out(fileName + ".vtk", std::ios::out | std::ios::binary);
    if (!out.is_open()) throw NeonException("ioToVtk: file " + fileName + ".vtk could not be opened");
    out << "# vtk DataFile Version 3.0" << std::endl;
    out << "Title Neon" << std::endl;
    out << (type == IoFileType::ASCII ? "ASCII" : "BINARY") << std::endl;
    out << "DATASET STRUCTURED_POINTS" << std::endl;
    out << "DIMENSIONS " << node.x << " " << node.y << " " << node.z << "\n";
    out << "ORIGIN " << 0.0 << " " << 0.0 << " " << 0.0 << "\n";
    out << "SPACING " << 1.0 << " " << 1.0 << " " << 1.0 << "\n";
    out << "CELL_DATA " << vox.rMulTyped() << std::endl;
    if (!isNodeSpace) writeScalars(out, fieldName, values, dim, card, type);
    if (includeDomain) writeScalars(out, "Domain", std::vector<V>(vox.rMulTyped(), V(0)), vox, 1, type);
    out << "POINT_DATA " << node.rMulTyped() << std::endl;
    if (isNodeSpace) writeScalars(out, fieldName, values, dim, card, type);
    if (!out.good()) throw NeonException("ioToVtk: writing " + fileName + ".vtk failed");
}

}  // namespace vtk

// XML ImageData (.vti) export, Neon::ioToVTI
// (libNeonCore/include/Neon/core/tools/io/ioToVti.h:33-44, :663-719).  Each
// input is a callback (index, component) -> value evaluated over the node
// box (isNode) or the voxel box, written as one Float64 DataArray: the
// PointData section first, then CellData (:468-539).  ASCII arrays print
// every value in fixed notation with 17 decimals followed by a space, a tab
// after each voxel's components, a newline after each x row and after each
// z plane (:431-462).  BINARY arrays hold the values as little-endian
// binary64 behind a UInt32 byte count, base64-encoded as one stream (the
// VTK XML inline-binary format); the reference's binary branch converts the
// values to its index type and sizes the buffer in elements (:374-415), so
// it is not reproduced.
struct double_3d
{
    double x = 0, y = 0, z = 0;
    double_3d() = default;
    double_3d(double x_, double y_, double z_) : x(x_), y(y_), z(z_) {}
};

template <class real_tt = double>
struct VtiInputData_t
{
    std::function<real_tt(const index_3d&, int componentIdx)> func;
    int                                                      nComponents = 1;
    std::string                                              fieldName;
    bool                                                     isNode = false;
    IoFileType                                               asciiOrBinary = IoFileType::ASCII;
};

namespace vti {

inline std::string base64(const std::vector<unsigned char>& in)
{
    static const char tab[] = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";
    std::string out;
    out.reserve((in.size() + 2) / 3 * 4);
    size_t i = 0;
    for (; i + 3 <= in.size(); i += 3) {
        const uint32_t w = (uint32_t(in[i]) << 16) | (uint32_t(in[i + 1]) << 8) | uint32_t(in[i + 2]);
        for (int s = 18; s >= 0; s -= 6) out += tab[(w >> s) & 63u];
    }
    const size_t rest = in.size() - i;
    if (rest) {
        uint32_t w = uint32_t(in[i]) << 16;
        if (rest == 2) w |= uint32_t(in[i + 1]) << 8;
        out += tab[(w >> 18) & 63u];
        out += tab[(w >> 12) & 63u];
        out += rest == 2 ? tab[(w >> 6) & 63u] : '=';
        out += '=';
    }
    return out;
}

inline std::string extent(const index_3d& box)
{
    return "0 " + std::to_string(box.x) + " 0 " + std::to_string(box.y) + " 0 " + std::to_string(box.z);
}

template <class real_tt>
inline void dataArray(std::ostream& out, const VtiInputData_t<real_tt>& in, const index_3d& box)
{
    out << "<DataArray type=\"Float64\" NumberOfComponents=\"" << in.nComponents << "\" Name=\"" << in.fieldName
        << "\" format=\"" << (in.asciiOrBinary == IoFileType::ASCII ? "ascii" : "binary") << "\">\n";
    if (in.asciiOrBinary == IoFileType::ASCII) {
        char buf[64];
        for (int z = 0; z < box.z; z++) {
            for (int y = 0; y < box.y; y++) {
                for (int x = 0; x < box.x; x++) {
                    for (int c = 0; c < in.nComponents; c++) {
                        const real_tt v = in.func(index_3d(x, y, z), c);
                        if constexpr (std::is_integral_v<real_tt>) out << static_cast<long long>(v) << " ";
                        else {
                            std::snprintf(buf, sizeof(buf), "%.17f ", static_cast<double>(v));
                            out << buf;
                        }
                    }
                    out << "\t";
                }
                out << "\n";
            }
            out << "\n";
        }
        out << "</DataArray>\n\n";
        return;
    }
    const size_t               n = box.rMulTyped() * size_t(in.nComponents);
    const size_t               bytes = n * sizeof(double);
    std::vector<unsigned char> raw(4 + bytes);
    const uint32_t             len = uint32_t(bytes);
    for (int k = 0; k < 4; k++) raw[size_t(k)] = static_cast<unsigned char>(len >> (8 * k));
    size_t at = 4;
    for (int z = 0; z < box.z; z++)
        for (int y = 0; y < box.y; y++)
            for (int x = 0; x < box.x; x++)
                for (int c = 0; c < in.nComponents; c++) {
                    const double v = static_cast<double>(in.func(index_3d(x, y, z), c));
                    uint64_t     u;
                    std::memcpy(&u, &v, sizeof(u));
                    for (int k = 0; k < 8; k++) raw[at++] = static_cast<unsigned char>(u >> (8 * k));
                }
    out << base64(raw) << "\n</DataArray>\n\n";
}

}  // namespace vti

// Writes `filename` (taken as given, no suffix added) for the inputs over a
// node box and a voxel box that must differ by one in every direction.
template <class real_tt = double>
inline void ioToVTI(const std::vector<VtiInputData_t<real_tt>>& gridsInfo, const std::string& filename,
                    const index_3d& nodeSpace, const index_3d& voxSpace,
                    const double_3d& spacingData = double_3d(1, 1, 1), const double_3d& origin = double_3d(0, 0, 0))
{
    if (voxSpace.x != nodeSpace.x - 1 || voxSpace.y != nodeSpace.y - 1 || voxSpace.z != nodeSpace.z - 1)
        throw NeonException("ioToVTI: Inconsistent data. Node space dimension should be bigger than voxel of one unit.");
    if (voxSpace.x < 0 || voxSpace.y < 0 || voxSpace.z < 0) throw NeonException("ioToVTI: negative voxel box");
    for (const auto& g : gridsInfo) {
        const index_3d box = g.isNode ? nodeSpace : voxSpace;
        if (g.asciiOrBinary == IoFileType::BINARY && box.rMulTyped() * size_t(g.nComponents) * 8 > 0xffffffffull)
            throw NeonException("ioToVTI: binary array " + g.fieldName + " exceeds the 4 GiB UInt32 header");
    }
    std::ofstream out(filename, std::ios::out | std::ios::binary);
    if (!out.is_open()) throw NeonException("ioToVTI: file " + filename + " could not be opened");
    const auto d = [](double v) { return std::to_string(v); };
    out << "<?xml version=\"1.0\"?>\n";
    out << "<VTKFile type=\"ImageData\" version=\"0.1\" byte_order=\"LittleEndian\">\n";
    out << "<ImageData WholeExtent=\"" << vti::extent(voxSpace) << " \" Origin=\"" << d(origin.x) << " "
        << d(origin.y) << " " << d(origin.z) << "\" Spacing=\"" << d(spacingData.x) << " " << d(spacingData.y)
        << " " << d(spacingData.z) << " \">\n";
    out << "<Piece Extent=\"" << vti::extent(voxSpace) << "\" >\n";
    for (const bool nodes : {true, false}) {
        out << (nodes ? "<PointData>\n" : "<CellData>\n");
        for (const auto& g : gridsInfo)
            if (g.isNode == nodes) vti::dataArray(out, g, nodes ? nodeSpace : voxSpace);
        out << (nodes ? "</PointData>\n" : "</CellData>\n");
    }
    out << "</Piece>\n";
    out << " </ImageData>\n";
    out << " </VTKFile>\n";
    if (!out.good()) throw NeonException("ioToVTI: writing " + filename + " failed");
}

// ---- Backend ---------------------------------------------------------------

class Backend
{
    struct State
    {
        int                    device = 0;
        std::vector<nx_stream> streams;
        ~State()
        {
            for (nx_stream s : streams) nx_stream_destroy(s);
        }
    };
    std::shared_ptr<State> s_;

   public:
    static constexpr int mainStreamIdx = 0;

    Backend() = default;
    explicit Backend(const std::vector<int>& devIds, Runtime = Runtime::stream, int nStreams = 1)
    {
        if (devIds.size() != 1)
            throw NeonException("Backend: one process drives one GPU; pass exactly one device id");
        check(nx_init(devIds[0]), "Backend");
        s_ = std::make_shared<State>();
        s_->device = devIds[0];
        for (int i = 0; i < nStreams; i++) {
            nx_stream st = nullptr;
            check(nx_stream_create(&st), "Backend: stream");
            s_->streams.push_back(st);
        }
    }
    nx_stream stream(int idx = mainStreamIdx) const
    {
        if (!s_ || idx < 0 || idx >= int(s_->streams.size())) throw NeonException("Backend: bad stream index");
        return s_->streams[idx];
    }
    void sync(int idx) const { check(nx_stream_sync(stream(idx)), "Backend::sync"); }
    void syncAll() const { check(nx_sync(), "Backend::syncAll"); }
    int  deviceId() const { return s_ ? s_->device : -1; }
    std::string toString() const
    {
        char name[256] = {0};
        nx_device_name(name, sizeof(name));
        return std::string("Backend gpu:") + std::to_string(deviceId()) + " " + name;
    }
};

// ---- cell classification field element --------------------------------------

// CellType (benchmarks/lbm-lid-driven-cavity-flow/src/CellType.h): packed on
// the device into one 32-bit word, classification in bits 24..25, the wall
// neighbour bit flags (and the library's flag bits 28, 29) in the rest.
struct CellType
{
    enum Classification : uint32_t { bounceBack = 0, movingWall = 1, bulk = 2, undefined = 3 };
    Classification classification = undefined;
    uint32_t       wallNghBitflag = 0;

    uint32_t pack() const
    {
        return (uint32_t(classification) << NX_CLASS_SHIFT) | (wallNghBitflag & ~(3u << NX_CLASS_SHIFT));
    }
    static CellType unpack(uint32_t w)
    {
        CellType c;
        c.classification = Classification((w >> NX_CLASS_SHIFT) & 3u);
        c.wallNghBitflag = w & ~(3u << NX_CLASS_SHIFT);
        return c;
    }
};

// Deviation storage types (populations stored as f - w_q; no counterpart in
// the reference): host mirrors hold the stored encoding.
struct F32Dev { float v = 0.f; };
struct F16Dev { uint16_t bits = 0; };

template <typename T> struct StorageOf;
template <> struct StorageOf<double> { static constexpr int value = NX_F64; };
template <> struct StorageOf<float> { static constexpr int value = NX_F32; };
template <> struct StorageOf<F32Dev> { static constexpr int value = NX_F32DEV; };
template <> struct StorageOf<F16Dev> { static constexpr int value = NX_F16DEV; };

// Compute types: double; float with the reference's expression semantics
// (its double literals promote parts of the float collision); FastFloat is
// float throughout (the bandwidth-bound fast path).
struct FastFloat {};
template <typename T> struct ComputeOf;
template <> struct ComputeOf<double> { static constexpr int value = NX_COMPUTE_F64; };
template <> struct ComputeOf<float> { static constexpr int value = NX_COMPUTE_REF; };
template <> struct ComputeOf<FastFloat> { static constexpr int value = NX_COMPUTE_F32; };

template <typename Grid, typename T, int C>
class Field;
class Container;

// ---- dGrid -------------------------------------------------------------------

// Dense grid of one partition (dGrid_imp.h:7-185): every cell of the box is
// allocated; field layout SoA, slot = card * n + (z * Y + y) * X + x.
class dGrid
{
    struct State
    {
        Backend  bk;
        index_3d dim;
        int      periodic = 0;
    };
    std::shared_ptr<State> s_;

   public:
    template <typename T, int C>
    using Field = neon_amd::Field<dGrid, T, C>;

    dGrid() = default;
    // The active-cell lambda is accepted and not used: the reference's dGrid
    // ignores it as well (every cell of the box is allocated and active,
    // dGrid_imp.h:7-10).  The stencil only sizes halos, which one partition
    // does not have.
    dGrid(const Backend& bk, index_3d dim, const std::function<bool(const index_3d&)>& = {},
          const void* /*stencil*/ = nullptr, int periodic = 0)
    {
        if (dim.x < 1 || dim.y < 1 || dim.z < 1) throw NeonException("dGrid: empty dimension");
        s_ = std::make_shared<State>();
        s_->bk = bk;
        s_->dim = dim;
        s_->periodic = periodic;
    }
    template <typename T, int C = 0>
    Field<T, C> newField(const std::string& name, int cardinality, T inactiveValue) const
    {
        return Field<T, C>(*this, name, cardinality, inactiveValue);
    }
    index_3d       getDimension() const { return s_->dim; }
    const Backend& getBackend() const { return s_->bk; }
    int            periodic() const { return s_->periodic; }
    size_t         nSlots() const { return s_->dim.rMulTyped(); }
    size_t         slot(size_t cell, int card, int) const { return size_t(card) * nSlots() + cell; }
    template <class F>
    void forEachActive(F&& f) const  // f(index_3d, cell)
    {
        const index_3d d = s_->dim;
        size_t         cell = 0;
        for (int z = 0; z < d.z; z++)
            for (int y = 0; y < d.y; y++)
                for (int x = 0; x < d.x; x++) f(index_3d(x, y, z), cell++);
    }
#if defined(__HIPCC__)
    // Neon's grid.newContainer(name, loadingLambda) (dGrid_imp.h:218-251):
    // the loading lambda receives a Loader, loads the fields it touches and
    // returns the compute lambda, `[=] __device__ (const dIndex& idx) mutable
    // {...}` (NEON_CUDA_HOST_DEVICE in the reference), which the container
    // launches once per cell of the box on run().  Compiled only when the
    // caller's translation unit is compiled with hipcc (defined below).
    template <typename LoadingLambda>
    Container newContainer(const std::string& name, LoadingLambda&& loadingLambda) const;
#endif
    nx_dense_desc desc(int storage, int compute) const
    {
        nx_dense_desc d{};
        d.dim[0] = s_->dim.x;
        d.dim[1] = s_->dim.y;
        d.dim[2] = s_->dim.z;
        d.storage = storage;
        d.compute = compute;
        d.periodic = s_->periodic;
        return d;
    }
};

// ---- bGrid -------------------------------------------------------------------

// Block-sparse grid of one partition, 8^3 blocks (bGrid_imp.h:19-246): blocks
// holding an active voxel; field layout [block][card][512], voxel x+8y+64z.
class bGrid
{
    struct State
    {
        Backend               bk;
        index_3d              dim;
        int                   periodic = 0;
        nx_bgrid              g = nullptr;
        uint32_t              nBlocks = 0;
        uint64_t              nActive = 0;
        std::vector<int32_t>  origins;
        std::vector<uint32_t> masks, conn;
        void *                dOrigins = nullptr, *dMasks = nullptr, *dConn = nullptr;
        // set by setStagedNeighbourhood(true); containers hold their own reference
        std::shared_ptr<nx_block_tiles_s> tiles;
        ~State()
        {
            tiles.reset();
            nx_free(dOrigins);
            nx_free(dMasks);
            nx_free(dConn);
            nx_bgrid_destroy(g);
        }
    };
    std::shared_ptr<State> s_;

    static int trampoline(int32_t x, int32_t y, int32_t z, void* user)
    {
        return (*static_cast<const std::function<bool(const index_3d&)>*>(user))(index_3d(x, y, z)) ? 1 : 0;
    }

   public:
    template <typename T, int C>
    using Field = neon_amd::Field<bGrid, T, C>;
    static constexpr int kBlockVoxels = 512;

    bGrid() = default;
    bGrid(const Backend& bk, index_3d dim, std::function<bool(const index_3d&)> activeCell,
          const void* /*stencil*/ = nullptr, int periodic = 0)
    {
        s_ = std::make_shared<State>();
        s_->bk = bk;
        s_->dim = dim;
        s_->periodic = periodic;
        if (!activeCell) activeCell = [](const index_3d&) { return true; };
        const int32_t d[3] = {dim.x, dim.y, dim.z};
        check(nx_bgrid_create_cb(d, &trampoline, &activeCell, periodic, &s_->g), "bGrid");
        check(nx_bgrid_info(s_->g, &s_->nBlocks, &s_->nActive), "bGrid");
        const size_t n = s_->nBlocks;
        s_->origins.resize(3 * n);
        s_->masks.resize(16 * n);
        s_->conn.resize(27 * n);
        check(nx_bgrid_tables(s_->g, s_->origins.data(), s_->masks.data(), s_->conn.data()), "bGrid");
        auto up = [&](const void* h, size_t bytes, void** dptr) {
            check(nx_alloc(bytes ? bytes : 4, dptr), "bGrid");
            if (bytes) check(nx_h2d(*dptr, h, bytes, nullptr), "bGrid");
        };
        up(s_->origins.data(), s_->origins.size() * 4, &s_->dOrigins);
        up(s_->masks.data(), s_->masks.size() * 4, &s_->dMasks);
        up(s_->conn.data(), s_->conn.size() * 4, &s_->dConn);
        check(nx_sync(), "bGrid");
    }
    template <typename T, int C = 0>
    Field<T, C> newField(const std::string& name, int cardinality, T inactiveValue) const
    {
        return Field<T, C>(*this, name, cardinality, inactiveValue);
    }
    index_3d       getDimension() const { return s_->dim; }
    const Backend& getBackend() const { return s_->bk; }
    uint32_t       nBlocks() const { return s_->nBlocks; }
    uint64_t       nActiveCells() const { return s_->nActive; }
    size_t         nSlots() const { return size_t(s_->nBlocks) * kBlockVoxels; }
    size_t         slot(size_t cell, int card, int cardinality) const
    {
        const size_t b = cell / kBlockVoxels, v = cell % kBlockVoxels;
        return (b * size_t(cardinality) + size_t(card)) * kBlockVoxels + v;
    }
    bool isActive(size_t cell) const
    {
        const size_t b = cell / kBlockVoxels, v = cell % kBlockVoxels;
        return (s_->masks[b * 16 + v / 32] >> (v % 32)) & 1u;
    }
    template <class F>
    void forEachActive(F&& f) const  // f(index_3d, cell = block * 512 + voxel)
    {
        for (uint32_t b = 0; b < s_->nBlocks; b++)
            for (int v = 0; v < kBlockVoxels; v++) {
                if (!((s_->masks[size_t(b) * 16 + v / 32] >> (v % 32)) & 1u)) continue;
                const index_3d o(s_->origins[3 * b], s_->origins[3 * b + 1], s_->origins[3 * b + 2]);
                f(index_3d(o.x + (v & 7), o.y + ((v >> 3) & 7), o.z + (v >> 6)), size_t(b) * kBlockVoxels + v);
            }
    }
    // LbmContainers::iteration on this grid steps with the block
    // neighbourhood staged in LDS (nx_d3q19_step_block_tiled, 2- and 4-byte
    // storage; the same populations bit for bit), or, off, one workgroup
    // per block (the default).  Containers made afterwards follow it.
    void setStagedNeighbourhood(bool on)
    {
        if (on && !s_->tiles) {
            const nx_block_desc d = desc(NX_F32, NX_COMPUTE_F32);
            nx_block_tiles      t = nullptr;
            check(nx_block_tiles_create(&d, &t), "bGrid::setStagedNeighbourhood");
            s_->tiles = std::shared_ptr<nx_block_tiles_s>(t, [](nx_block_tiles_s* p) { nx_block_tiles_destroy(p); });
        } else if (!on) {
            s_->tiles.reset();
        }
    }
    std::shared_ptr<nx_block_tiles_s> stagedTiles() const { return s_->tiles; }
    nx_block_desc desc(int storage, int compute) const
    {
        nx_block_desc d{};
        d.dim[0] = s_->dim.x;
        d.dim[1] = s_->dim.y;
        d.dim[2] = s_->dim.z;
        d.storage = storage;
        d.compute = compute;
        d.periodic = s_->periodic;
        d.n_blocks = s_->nBlocks;
        d.origins = static_cast<const int32_t*>(s_->dOrigins);
        d.masks = static_cast<const uint32_t*>(s_->dMasks);
        d.conn = static_cast<const uint32_t*>(s_->dConn);
        return d;
    }
};

// ---- Field -------------------------------------------------------------------

template <typename T> struct DeviceRepr { using type = T; };
template <> struct DeviceRepr<CellType> { using type = uint32_t; };

// A field of `cardinality` values per cell with a host mirror (the
// reference's MemMirror): forEachActiveCell edits the host copy,
// updateDeviceData / updateHostData move it (asynchronously on the stream).
template <typename GridT, typename T, int C>
class Field
{
    using D = typename DeviceRepr<T>::type;
    struct State
    {
        GridT          grid;
        std::string    name;
        int            card = 1;
        T              inactive{};
        std::vector<T> host;
        std::vector<D> staging;
        void*          dev = nullptr;
        ~State() { nx_free(dev); }
    };
    std::shared_ptr<State> s_;

   public:
    using Type = T;
    using Grid = GridT;

    Field() = default;
    Field(const GridT& g, const std::string& name, int cardinality, T inactiveValue)
    {
        if (C != 0 && cardinality != C) throw NeonException("newField: cardinality differs from the template's");
        s_ = std::make_shared<State>();
        s_->grid = g;
        s_->name = name;
        s_->card = cardinality;
        s_->inactive = inactiveValue;
        const size_t n = g.nSlots() * size_t(cardinality);
        s_->host.assign(n, inactiveValue);
        check(nx_alloc(n * sizeof(D), &s_->dev), "newField " + name);
        updateDeviceData(Backend::mainStreamIdx);
        g.getBackend().sync(Backend::mainStreamIdx);
    }
#if defined(__HIPCC__)
    // The device view a container's compute lambda captures (Neon's
    // field.getPartition(execution, setIdx, dataView), dField_imp.h:67-87):
    // raw pointer, dimensions, cardinality; valid while the field lives.
    auto getPartition() const;
#endif
    bool           valid() const { return bool(s_); }
    const GridT&   getGrid() const { return s_->grid; }
    const Backend& getBackend() const { return s_->grid.getBackend(); }
    int            getCardinality() const { return s_->card; }
    const std::string& getName() const { return s_->name; }
    void*          devPtr() const { return s_->dev; }

    template <class F>
    void forEachActiveCell(F&& f)  // f(const index_3d&, const int& card, T& value)
    {
        const int card = s_->card;
        s_->grid.forEachActive([&](const index_3d& idx, size_t cell) {
            for (int k = 0; k < card; k++) f(idx, k, s_->host[s_->grid.slot(cell, k, card)]);
        });
    }
    T& hostRef(size_t cell, int card) { return s_->host[s_->grid.slot(cell, card, s_->card)]; }

    // Writes fileName + ".vtk" from the host mirror (call updateHostData and
    // sync first, as the reference's callers do), FieldBase::ioToVtk
    // (libNeonDomain/include/Neon/domain/interface/FieldBase_imp.h:297-318,
    // IOGridVTK_imp.h:6-64): every cell of the grid box, inactive cells
    // holding the field's inactive value; includeDomain adds the cell-data
    // field "Domain" with the owning partition (0: one partition).
    template <typename VtiExportType = T>
    void ioToVtk(const std::string& fileName, const std::string& fieldName, bool includeDomain = false,
                 IoFileType ioFileType = IoFileType::ASCII, bool isNodeSpace = false) const
    {
        const index_3d dim = s_->grid.getDimension();
        const int      card = s_->card;
        std::vector<VtiExportType> dense(dim.rMulTyped() * size_t(card), VtiExportType(s_->inactive));
        s_->grid.forEachActive([&](const index_3d& idx, size_t cell) {
            const size_t at = (size_t(idx.z) * size_t(dim.y) + size_t(idx.y)) * size_t(dim.x) + size_t(idx.x);
            for (int k = 0; k < card; k++)
                dense[at * size_t(card) + size_t(k)] = VtiExportType(s_->host[s_->grid.slot(cell, k, card)]);
        });
        vtk::writeFile(fileName, fieldName, dense, dim, card, includeDomain, ioFileType, isNodeSpace);
    }

    void updateDeviceData(int streamIdx)
    {
        const size_t n = s_->host.size();
        const void*  src = s_->host.data();
        if constexpr (!std::is_same_v<T, D>) {
            s_->staging.resize(n);
            for (size_t i = 0; i < n; i++) s_->staging[i] = s_->host[i].pack();
            src = s_->staging.data();
        }
        check(nx_h2d(s_->dev, src, n * sizeof(D), getBackend().stream(streamIdx)), "updateDeviceData " + s_->name);
    }
    void updateHostData(int streamIdx)
    {
        const size_t n = s_->host.size();
        if constexpr (std::is_same_v<T, D>) {
            check(nx_d2h(s_->host.data(), s_->dev, n * sizeof(D), getBackend().stream(streamIdx)),
                  "updateHostData " + s_->name);
        } else {
            s_->staging.resize(n);
            check(nx_d2h(s_->staging.data(), s_->dev, n * sizeof(D), getBackend().stream(streamIdx)),
                  "updateHostData " + s_->name);
            getBackend().sync(streamIdx);
            for (size_t i = 0; i < n; i++) s_->host[i] = T::unpack(s_->staging[i]);
        }
    }
};

// ---- Container / Skeleton --------------------------------------------------

class Container
{
    std::string                     name_;
    Backend                         bk_;
    std::function<void(nx_stream)>  body_;
    std::shared_ptr<void>           keep_;  // resources the body uses

   public:
    Container() = default;
    Container(std::string name, const Backend& bk, std::function<void(nx_stream)> body,
              std::shared_ptr<void> keep = {})
        : name_(std::move(name)), bk_(bk), body_(std::move(body)), keep_(std::move(keep)) {}
    void run(int streamIdx = Backend::mainStreamIdx) const { body_(bk_.stream(streamIdx)); }
    const std::string& getName() const { return name_; }
};

struct SkeletonOptions
{
    Occ          occ = Occ::none;
    TransferMode transfer = TransferMode::get;
    SkeletonOptions() = default;
    SkeletonOptions(Occ o, TransferMode t) : occ(o), transfer(t) {}
};

// A sequence of containers run in order on the main stream (with one
// partition there is no halo update to insert between them).  The
// SkeletonOptions (OCC, transfer mode) are accepted and have no effect: they
// choose how halo transfers overlap computation across partitions
// (Skeleton.h:32-65, skeleton/Options.h), and one partition has none.
class Skeleton
{
    Backend                bk_;
    std::vector<Container> seq_;
    std::string            name_;

   public:
    Skeleton() = default;
    explicit Skeleton(const Backend& bk) : bk_(bk) {}
    void sequence(const std::vector<Container>& ops, const std::string& name, SkeletonOptions = {})
    {
        seq_ = ops;
        name_ = name;
    }
    void run() const
    {
        for (const Container& c : seq_) c.run(Backend::mainStreamIdx);
    }
    const std::string& getName() const { return name_; }
};

// ---- D3Q19 lattice-Boltzmann ------------------------------------------------

namespace lbm {

// D3Q19 velocities in the reference's order (D3Q19.h:22-42): opposite of q is
// q +- 10 (9 is the rest direction); weights 1/3, 1/18, 1/36.
struct D3Q19
{
    static constexpr int Q = 19;
    static constexpr int centerDirection = 9;
    static constexpr int goBackOffset = 10;
    std::vector<index_3d> c_vect;
    std::vector<double>   t_vect;
    D3Q19()
    {
        const int cx[Q] = {-1, 0, 0, -1, -1, -1, -1, 0, 0, 0, 1, 0, 0, 1, 1, 1, 1, 0, 0};
        const int cy[Q] = {0, -1, 0, -1, 1, 0, 0, -1, -1, 0, 0, 1, 0, 1, -1, 0, 0, 1, 1};
        const int cz[Q] = {0, 0, -1, 0, 0, -1, 1, -1, 1, 0, 0, 0, 1, 0, 0, 1, -1, 1, -1};
        for (int q = 0; q < Q; q++) {
            c_vect.emplace_back(cx[q], cy[q], cz[q]);
            const int l1 = std::abs(cx[q]) + std::abs(cy[q]) + std::abs(cz[q]);
            t_vect.push_back(l1 == 0 ? 1. / 3. : (l1 == 1 ? 1. / 18. : 1. / 36.));
        }
    }
};

namespace detail {
template <typename G> struct IsDense : std::false_type {};
template <> struct IsDense<dGrid> : std::true_type {};
}  // namespace detail

// LbmContainers<Lattice, PopulationField, LbmComputeType> (LbmTools.h:22-437):
// the step, the wall-neighbour mask and the moments as containers.
template <typename PopulationField, typename ComputeType>
struct LbmContainers
{
    using Grid = typename PopulationField::Grid;
    using Store = typename PopulationField::Type;
    using CellTypeField = Field<Grid, CellType, 1>;
    static constexpr int kStorage = StorageOf<Store>::value;
    static constexpr int kCompute = ComputeOf<ComputeType>::value;

    template <typename Omega>
    static Container iteration(StencilSemantic, const PopulationField& fIn, const CellTypeField& cellTypeField,
                               Omega omega, PopulationField& fOut, const double force[3] = nullptr)
    {
        const Grid&   g = fIn.getGrid();
        nx_lbm_params p{};
        p.omega = double(omega);
        for (int a = 0; a < 3; a++) p.force[a] = force ? force[a] : 0.0;
        const void*     fin = fIn.devPtr();
        void*           fout = fOut.devPtr();
        const uint32_t* flags = static_cast<const uint32_t*>(cellTypeField.devPtr());
        if constexpr (detail::IsDense<Grid>::value) {
            // The library container classifies the cells when it is created,
            // so it is bound at the first run: the reference builds its
            // iteration before the cell types are set (RunCavityTwoPop.cu:77-84
            // vs 208-241).  Cell types and wall values must not change after
            // that run (the container notes the walls it finds).
            struct Lazy
            {
                nx_dense_desc   d;
                const void*     fin;
                void*           fout;
                const uint32_t* flags;
                nx_lbm_params   p;
                nx_container    c = nullptr;
                ~Lazy() { nx_container_destroy(c); }
            };
            auto lazy = std::make_shared<Lazy>(Lazy{g.desc(kStorage, kCompute), fin, fout, flags, p});
            return Container("LbmIteration", g.getBackend(),
                             [lazy](nx_stream s) {
                                 if (!lazy->c) {
                                     check(nx_stream_sync(s), "LbmIteration");
                                     check(nx_d3q19_iteration_dense(&lazy->d, lazy->fin, lazy->fout, lazy->flags,
                                                                    &lazy->p, &lazy->c),
                                           "LbmIteration");
                                 }
                                 check(nx_container_run(lazy->c, s), "LbmIteration");
                             },
                             lazy);
        } else {
            const nx_block_desc                     d = g.desc(kStorage, kCompute);
            const std::shared_ptr<nx_block_tiles_s> t = g.stagedTiles();
            return Container("LbmIteration", g.getBackend(), [d, t, fin, fout, flags, p](nx_stream s) {
                if (t)
                    check(nx_d3q19_step_block_tiled(&d, t.get(), fin, fout, flags, &p, s), "LbmIteration");
                else
                    check(nx_d3q19_step_block(&d, fin, fout, flags, &p, s), "LbmIteration");
            });
        }
    }

    static Container computeWallNghMask(const CellTypeField& infoInField, CellTypeField& infoOutField)
    {
        const Grid&     g = infoInField.getGrid();
        const uint32_t* in = static_cast<const uint32_t*>(infoInField.devPtr());
        uint32_t*       out = static_cast<uint32_t*>(infoOutField.devPtr());
        if constexpr (detail::IsDense<Grid>::value) {
            const nx_dense_desc d = g.desc(kStorage, kCompute);
            return Container("LBM_computeWallNghMask", g.getBackend(), [d, in, out](nx_stream s) {
                check(nx_d3q19_wall_mask_dense(&d, in, out, s), "computeWallNghMask");
            });
        } else {
            const nx_block_desc d = g.desc(kStorage, kCompute);
            return Container("LBM_computeWallNghMask", g.getBackend(), [d, in, out](nx_stream s) {
                check(nx_d3q19_wall_mask_block(&d, in, out, s), "computeWallNghMask");
            });
        }
    }

    // rho [1], u [3]: double fields for double storage, float otherwise
    template <typename RhoField, typename UField>
    static Container computeRhoAndU(const PopulationField& fInField, const CellTypeField& cellTypeField,
                                    RhoField& rhoField, UField& uField)
    {
        using M = std::conditional_t<kStorage == NX_F64, double, float>;
        static_assert(std::is_same_v<typename RhoField::Type, M> && std::is_same_v<typename UField::Type, M>,
                      "rho and u fields: double for double storage, float otherwise");
        const Grid&     g = fInField.getGrid();
        const void*     fin = fInField.devPtr();
        const uint32_t* flags = static_cast<const uint32_t*>(cellTypeField.devPtr());
        void*           rho = rhoField.devPtr();
        void*           u = uField.devPtr();
        if constexpr (detail::IsDense<Grid>::value) {
            const nx_dense_desc d = g.desc(kStorage, kCompute);
            return Container("LBM_computeRhoAndU", g.getBackend(), [d, fin, flags, rho, u](nx_stream s) {
                check(nx_d3q19_moments_dense(&d, fin, flags, rho, u, s), "computeRhoAndU");
            });
        } else {
            const nx_block_desc d = g.desc(kStorage, kCompute);
            return Container("LBM_computeRhoAndU", g.getBackend(), [d, fin, flags, rho, u](nx_stream s) {
                check(nx_d3q19_moments_block(&d, fin, flags, rho, u, s), "computeRhoAndU");
            });
        }
    }
};

// LbmIterationD3Q19 (LbmIteration.h:17-101): two skeletons, one per
// direction of the population ping-pong, alternated by run().
template <typename PopulationField, typename ComputeType>
struct LbmIterationD3Q19
{
    using Tools = LbmContainers<PopulationField, ComputeType>;
    using CellTypeField = typename Tools::CellTypeField;

    template <typename Omega>
    LbmIterationD3Q19(StencilSemantic stencilSemantic, Occ occ, TransferMode transfer, PopulationField& fIn,
                      PopulationField& fOut, CellTypeField& cellTypeField, Omega omega,
                      const double force[3] = nullptr)
    {
        pop[0] = fIn;
        pop[1] = fOut;
        for (int t = 0; t < 2; t++) {
            lbmTwoPop[t] = Skeleton(fIn.getBackend());
            lbmTwoPop[t].sequence({Tools::iteration(stencilSemantic, pop[t], cellTypeField, omega, pop[1 - t], force)},
                                  "LBM_iteration_" + std::to_string(t), SkeletonOptions(occ, transfer));
        }
        parity = 0;
    }
    PopulationField& getInput() { return pop[parity]; }
    PopulationField& getOutput() { return pop[parity ^ 1]; }
    void             run()
    {
        lbmTwoPop[parity].run();
        parity ^= 1;
    }
    void sync() { pop[0].getBackend().syncAll(); }

   private:
    Skeleton        lbmTwoPop[2];
    PopulationField pop[2];
    int             parity = 0;
};

}  // namespace lbm

// ---- multi-resolution grid with its LBM ---------------------------------------

// Neon's mGrid (libNeonDomain/include/Neon/domain/details/mGrid/mGrid.h:23-60)
// together with the lbmMultiRes application's LBM on it: one activation
// lambda per level (evaluated at each level-l voxel's level-0 origin
// v 2^l, as the reference's constructor does, mGrid.cpp:89-156), depth =
// number of lambdas (2..5), fields of T = double or float, the app's KBC
// D3Q27 (its default build) or BGK D3Q19 collision.  initLidDrivenCavity
// and run replace the app's initLidDrivenCavity + initSumStore
// (lidDrivenCavity.h:8-78, init.h:8-86) and the skeleton of
// nonUniformTimestepRecursive (lbmMultiRes.h:141-263, 300-376).
class mGrid
{
    struct State
    {
        Backend  bk;
        nx_mgrid g = nullptr;
        index_3d dim;
        int      depth = 0, storage = NX_F64, q = 19;
        ~State()
        {
            if (g) nx_mgrid_destroy(g);
        }
    };
    std::shared_ptr<State> s_;

   public:
    enum class Collision { BGK, KBC };  // the app's #define BGK / #define KBC

    mGrid() = default;
    template <typename T>
    static mGrid make(const Backend& bk, const index_3d& dim,
                      const std::vector<std::function<bool(const index_3d&)>>& activeCellLambda,
                      Collision collision = Collision::KBC)
    {
        static_assert(std::is_same_v<T, double> || std::is_same_v<T, float>, "mGrid fields are double or float");
        mGrid m;
        m.s_ = std::make_shared<State>();
        State& s = *m.s_;
        s.bk = bk;
        s.dim = dim;
        s.depth = int(activeCellLambda.size());
        s.storage = std::is_same_v<T, double> ? NX_F64 : NX_F32;
        std::vector<std::vector<uint8_t>> lam(s.depth);
        std::vector<const uint8_t*>       ptr(s.depth);
        index_3d                          d = dim;
        for (int l = 0; l < s.depth; l++) {
            lam[l].resize(d.rMulTyped());
            size_t i = 0;
            for (int z = 0; z < d.z; z++)
                for (int y = 0; y < d.y; y++)
                    for (int x = 0; x < d.x; x++) lam[l][i++] = activeCellLambda[l](index_3d(x << l, y << l, z << l));
            ptr[l] = lam[l].data();
            d = index_3d((d.x + 1) / 2, (d.y + 1) / 2, (d.z + 1) / 2);
        }
        const int32_t dd[3] = {dim.x, dim.y, dim.z};
        check(nx_mgrid_create_levels(dd, s.depth, ptr.data(), s.storage,
                                     collision == Collision::KBC ? NX_MGRID_KBC27 : NX_MGRID_BGK19, &s.g),
              "mGrid");
        check(nx_mgrid_populations(s.g, &s.q), "mGrid");
        return m;
    }

    int      getDepth() const { return s_->depth; }
    int      populations() const { return s_->q; }
    index_3d getDimension() const { return s_->dim; }
    // voxels per axis of level l: ceil(dim / 2^l)
    index_3d levelDimension(int level) const
    {
        std::vector<int32_t> d(3 * size_t(s_->depth));
        check(nx_mgrid_depth(s_->g, nullptr, d.data()), "mGrid::levelDimension");
        return index_3d(d[3 * level], d[3 * level + 1], d[3 * level + 2]);
    }
    // mGrid::getDimension(level) (mGrid.cpp:684-687): the level's blocks of
    // 2^3 voxels times 2
    index_3d getDimension(int level) const
    {
        const int sp = 1 << (level + 1);
        return index_3d((s_->dim.x + sp - 1) / sp * 2, (s_->dim.y + sp - 1) / sp * 2, (s_->dim.z + sp - 1) / sp * 2);
    }
    // leaf voxels per level (the app's countActiveVoxels, init.h:93-130)
    std::vector<uint64_t> numLeafVoxels() const
    {
        std::vector<uint32_t> nb(size_t(s_->depth));
        std::vector<uint64_t> na(size_t(s_->depth)), nl(size_t(s_->depth));
        check(nx_mgrid_info(s_->g, nb.data(), na.data(), nl.data(), nullptr, nullptr), "mGrid::numLeafVoxels");
        return nl;
    }
    void initLidDrivenCavity(double ulb) { check(nx_mgrid_cavity(s_->g, ulb), "initLidDrivenCavity"); }
    // the set-up of any problem (flowOverShape.h:17-113): cls[l] holds the
    // class of every level-l voxel (NX_CELL_BULK, NX_CELL_BOUNCEBACK,
    // NX_CELL_MOVINGWALL for the lid or an inlet), uWall the velocity of the
    // moving-wall cells
    void initCells(const std::vector<std::vector<uint8_t>>& cls, const double uWall[3])
    {
        if (int(cls.size()) != s_->depth) throw NeonException("mGrid::initCells: one class array per level");
        std::vector<const uint8_t*> ptr;
        for (int l = 0; l < s_->depth; l++) {
            if (cls[size_t(l)].size() != levelDimension(l).rMulTyped())
                throw NeonException("mGrid::initCells: class array size differs from the level's");
            ptr.push_back(cls[size_t(l)].data());
        }
        check(nx_mgrid_setup_cells(s_->g, ptr.data(), uWall), "mGrid::initCells");
    }
    // nIter steps of the coarsest level with its relaxation omega0
    void run(double omega0, int nIter, int streamIdx = Backend::mainStreamIdx)
    {
        check(nx_mgrid_step(s_->g, omega0, nIter, s_->bk.stream(streamIdx)), "mGrid::run");
    }
    // state bytes of a level: bit 0 active, bit 1 refined, bits 2..3 the class
    std::vector<uint8_t> state(int level) const
    {
        std::vector<uint8_t> st(levelDimension(level).rMulTyped());
        check(nx_mgrid_state_level(s_->g, level, st.data()), "mGrid::state");
        return st;
    }
    // fin (fout = false) or fout of a level, dense [q][z][y][x] in T
    template <typename T>
    std::vector<T> populations(int level, bool fout) const
    {
        if ((s_->storage == NX_F64) != std::is_same_v<T, double>)
            throw NeonException("mGrid::populations: the field type differs");
        std::vector<T> f(levelDimension(level).rMulTyped() * size_t(s_->q));
        check(nx_mgrid_fields(s_->g, level, fout ? 1 : 0, f.data()), "mGrid::populations");
        return f;
    }
};
}  // namespace neon_amd

#if defined(__HIPCC__)
// ---- user containers on the dense grid (hipcc translation units only) -------
// The reference's Container API for a user's own stencil: Loader::load
// (Loader_imp.h:102-188) returns the field's partition; the partition offers
// operator()(idx, card), getNghData<dx, dy, dz>(idx, card[, alternative])
// and getNghData(idx, offset, card) returning NghData{mData, mIsValid}
// (dPartition.h:174-190, 265-299; NghData.h), getGlobalIndex and
// cardinality.  A neighbour outside the box is invalid unless the grid is
// periodic along that axis (then it wraps).
#include <hip/hip_runtime.h>

namespace neon_amd {

// A cell of the dense grid (dIndex.h): its coordinates in the box.
struct dIndex
{
    int32_t x, y, z;
};

struct int8_3d
{
    int8_t x = 0, y = 0, z = 0;
    __host__ __device__ int8_3d() = default;
    __host__ __device__ int8_3d(int8_t x_, int8_t y_, int8_t z_) : x(x_), y(y_), z(z_) {}
};

template <typename T>
struct NghData
{
    T    mData;
    bool mIsValid;
};

// Neon::Pattern: how a container reads a field (Loader_imp.h:102-188)
enum class Pattern { MAP, STENCIL };

template <typename T>
struct dPartition
{
    T*       mem = nullptr;
    int32_t  X = 0, Y = 0, Z = 0;
    int32_t  card = 1;
    int32_t  periodic = 0;
    uint64_t n = 0;  // cells

    __host__ __device__ uint64_t cell(const dIndex& i) const
    {
        return uint64_t(i.x) + uint64_t(X) * (uint64_t(i.y) + uint64_t(Y) * uint64_t(i.z));
    }
    __device__ T& operator()(const dIndex& i, int c) const { return mem[uint64_t(c) * n + cell(i)]; }
    __device__ int cardinality() const { return card; }
    __device__ index_3d getGlobalIndex(const dIndex& i) const { return index_3d(i.x, i.y, i.z); }
    __device__ bool nghIdx(const dIndex& i, int dx, int dy, int dz, dIndex& out) const
    {
        int x = i.x + dx, y = i.y + dy, z = i.z + dz;
        if (periodic & NX_PERIODIC_X) x = (x + X) % X;
        if (periodic & NX_PERIODIC_Y) y = (y + Y) % Y;
        if (periodic & NX_PERIODIC_Z) z = (z + Z) % Z;
        out = dIndex{x, y, z};
        return x >= 0 && y >= 0 && z >= 0 && x < X && y < Y && z < Z;
    }
    __device__ NghData<T> getNghData(const dIndex& i, int8_3d off, int c, T alternative = T()) const
    {
        dIndex j;
        if (!nghIdx(i, off.x, off.y, off.z, j)) return NghData<T>{alternative, false};
        return NghData<T>{(*this)(j, c), true};
    }
    template <int dx, int dy, int dz>
    __device__ NghData<T> getNghData(const dIndex& i, int c, T alternative = T()) const
    {
        return getNghData(i, int8_3d(dx, dy, dz), c, alternative);
    }
};

// Packed cell type word (CellType.h) as the partition of a CellType field
// holds it on the device.
__device__ inline int cellClassification(uint32_t w) { return int((w >> NX_CLASS_SHIFT) & 3u); }
__device__ inline uint32_t cellWallNghBitflag(uint32_t w) { return w & ~(3u << NX_CLASS_SHIFT); }

class Loader
{
   public:
    // A STENCIL load of a field the container also writes is refused, as the
    // reference refuses a stencil load of a non-const field
    // (Loader_imp.h:114-122): neighbours would be read while written.
    template <typename FieldT>
    auto load(FieldT& field, Pattern pattern = Pattern::MAP,
              StencilSemantic = StencilSemantic::standard) const
    {
        if (pattern == Pattern::STENCIL && !std::is_const_v<FieldT>)
            throw NeonException("Loader::load: STENCIL access needs a const field");
        return field.getPartition();
    }
};

template <typename GridT, typename T, int C>
auto Field<GridT, T, C>::getPartition() const
{
    static_assert(std::is_same_v<GridT, dGrid>, "getPartition: dense grid fields only");
    dPartition<D> p;
    const index_3d d = s_->grid.getDimension();
    p.mem = static_cast<D*>(s_->dev);
    p.X = d.x;
    p.Y = d.y;
    p.Z = d.z;
    p.card = s_->card;
    p.periodic = s_->grid.periodic();
    p.n = d.rMulTyped();
    return p;
}

namespace detail {
// One thread per cell of the box, 256 cells per workgroup along x-fastest
// cell order: each wavefront's accesses to a SoA plane are contiguous.
template <typename F>
__global__ void __launch_bounds__(256) dgrid_launch(F f, int32_t X, int32_t Y, uint64_t n)
{
    const uint64_t c = uint64_t(blockIdx.x) * 256u + threadIdx.x;
    if (c >= n) return;
    const dIndex i{int32_t(c % uint64_t(X)), int32_t((c / uint64_t(X)) % uint64_t(Y)),
                   int32_t(c / (uint64_t(X) * uint64_t(Y)))};
    f(i);
}
}  // namespace detail

// The loading lambda runs once, when the container is made (the reference
// runs it again at every launch to collect the partitions; the partitions
// here are fixed for the fields' lifetime).
template <typename LoadingLambda>
Container dGrid::newContainer(const std::string& name, LoadingLambda&& loadingLambda) const
{
    Loader     loader;
    auto       compute = loadingLambda(loader);
    const auto d = getDimension();
    return Container(name, getBackend(), [compute, d, name](nx_stream s) {
        const uint64_t n = d.rMulTyped();
        hipLaunchKernelGGL(detail::dgrid_launch<decltype(compute)>, dim3(unsigned((n + 255) / 256)), dim3(256), 0,
                           reinterpret_cast<hipStream_t>(s), compute, d.x, d.y, n);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) throw NeonException(name + ": " + hipGetErrorString(e));
    });
}

}  // namespace neon_amd
#endif  // __HIPCC__
