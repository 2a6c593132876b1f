// sassena_host.hpp -- host C++ side of the scattering-device hot path.
//
// Mirrors the reference's scatter_devices / stager / control interface
// (IScatterDevice, AllVectorsScatterDevice, SelfVectorsScatterDevice,
// ScatterDeviceFactory, DataStagerByFrame, DataStagerByAtom, ScatterFactors,
// the orientation-vector and q-scan generators) over the C ABI of
// libsassena_amd.so.  MPI partitions, worker threads and TCP services of the
// reference are replaced by one GPU context per device; all arithmetic of the
// path runs in the HIP library.
#pragma once

#include <complex>
#include <cstddef>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "sassena_amd.h"

namespace sassena_amd {

// ------------------------------------------------------------ basics ------

// CartesianCoor3D (include/math/coor3d.hpp:35-70): double x, y, z.
struct CartesianCoor3D {
    double x = 0, y = 0, z = 0;
    CartesianCoor3D() = default;
    CartesianCoor3D(double a, double b, double c) : x(a), y(b), z(c) {}
    double length() const;
    CartesianCoor3D operator+(const CartesianCoor3D &o) const { return {x + o.x, y + o.y, z + o.z}; }
    CartesianCoor3D operator-(const CartesianCoor3D &o) const { return {x - o.x, y - o.y, z - o.z}; }
    double dot(const CartesianCoor3D &o) const { return x * o.x + y * o.y + z * o.z; }
    CartesianCoor3D cross_product(const CartesianCoor3D &o) const;
};
CartesianCoor3D operator*(double s, const CartesianCoor3D &v);
CartesianCoor3D operator/(const CartesianCoor3D &v, double s);

// sassena::terminate_request (include/exceptions/exceptions.hpp): the typed
// error the reference raises when a run cannot proceed.
class terminate_request : public std::runtime_error {
   public:
    explicit terminate_request(const std::string &what) : std::runtime_error(what) {}
};

// A failure reported by the device library (negative SX_ERR_* code).
class device_error : public std::runtime_error {
   public:
    device_error(int code, const std::string &what) : std::runtime_error(what), code(code) {}
    int code;
};

// Named start/stop intervals, as the reference's Timer (report/timer.cpp:35-59).
class Timer {
   public:
    struct Stat {
        double sum = 0, min = 1e300, max = 0;
        size_t count = 0;
    };
    void start(const std::string &label);
    void stop(const std::string &label);
    // adds n to the count of `label` without a time (events, not intervals)
    void add_count(const std::string &label, size_t n);
    const std::map<std::string, Stat> &stats() const { return stats_; }

   private:
    std::map<std::string, double> open_;
    std::map<std::string, Stat> stats_;
};

// ------------------------------------------------------------ control -----

// ScatteringVectorsScanParameters (parameters.hpp:815-845)
struct ScanParameters {
    double from = 0, to = 0;
    size_t points = 100;
    double exponent = 1.0;
    CartesianCoor3D basevector{1, 0, 0};
};

// ScatteringAverageOrientationVectorsParameters (parameters.cpp:930-1034)
struct OrientationVectors {
    std::string type = "sphere";                        // sphere | cylinder | file
    std::string algorithm = "boost_uniform_on_sphere";  // or raster_linear (cylinder)
    long resolution = 100;
    unsigned long seed = 0;
    std::string filepath = "qvector-orientations.txt";
    std::vector<CartesianCoor3D> vectors;
    void create();
};

// ScatteringAverageOrientationMultipoleMomentsParameters (parameters.hpp:621-648,
// parameters.cpp:1037-1122): (major, minor) pairs read from a file or generated
// up to a maximum major, then checked against the multipole type.
struct MultipoleMoments {
    std::string type;  // file | resolution (no default, as the reference)
    long resolution = 20;
    std::string filepath = "moments.txt";
    std::vector<std::pair<long, long>> moments;
    void create(const std::string &multipole_type);
};

struct ScatteringParameters {
    std::string type = "all";              // all | self
    double background_factor = 0.0;
    std::string dsp_type = "autocorrelate";  // autocorrelate | square | plain
    std::string dsp_method = "fftw";         // fftw | direct
    std::string orientation_type = "none";   // none | vectors | multipole
    CartesianCoor3D axis{0, 0, 1};
    OrientationVectors vectors;
    std::string multipole_type = "sphere";  // sphere | cylinder
    MultipoleMoments moments;
    std::vector<CartesianCoor3D> qvectors;   // scattering.vectors (single | scans | file)
    std::vector<ScanParameters> scans;
    void create_from_scans();                // parameters.cpp:1125-1189
};

// limits.stage.memory.data and limits.computation.memory.{result_buffer,scale}
// with the reference's defaults (parameters.cpp:611-624).
struct MemoryLimits {
    size_t stage_data = (size_t)500 * 1024 * 1024;
    size_t result_buffer = (size_t)100 * 1024 * 1024;
    size_t scale = 1;
};

struct Params {
    ScatteringParameters scattering;
    std::string stager_target = "system";
    int device = 0;
    size_t batch_shells = 64;  // q shells per device launch sequence
    MemoryLimits limits;
};

// The memory checks that precede staging (AbstractScatterDevice::ram_check,
// abstract_scatter_device.cpp:80-103, called from run() at :111, and the
// stager's coordinate buffer check, data_stager.cpp:55-63): the result buffer
// NF x 16 bytes against scale x limits.computation.memory.result_buffer and
// the staged block NA x NF x 12 bytes against limits.stage.memory.data.
// Throws terminate_request naming the limit, its value and the request.
void ram_check(const Params &params, size_t NA, size_t NF);

// ------------------------------------------------------------ database ----

// One tabulated function of the reference database (database.cpp:391-528):
// function_type selects constant (0), Slater (1) or Cromer-Mann (2) forms.
struct DatabaseFunction {
    int function_type = 0;
    std::vector<double> constants;
};

struct ElementEntry {
    DatabaseFunction sfactor;    // DatabaseSFactorsParameters::get
    DatabaseFunction volume;     // DatabaseVolumesParameters::get
    DatabaseFunction exclusion;  // DatabaseExlusionParameters::get
};

class Database {
   public:
    void add(const std::string &element, const ElementEntry &e);
    size_t id(const std::string &element) const;
    double sfactor(size_t id, double q) const;
    double volume(size_t id) const;
    double exclusion(size_t id, double effvolume, double q) const;
    // Neutron coherent scattering lengths (fm) of common elements, constant
    // in q (function type 0).
    static Database neutron_coherent();
    // The entries as a database document in the layout Database::read_xml
    // reads (database.cpp:31-145): scatterfactors, sizes and
    // exclusionfactors, one <element> with <name>, <type> and <param>s each.
    std::string to_xml() const;

   private:
    std::vector<std::string> names_;
    std::vector<ElementEntry> entries_;
};

// ScatterFactors (scatter_factors.cpp:28-130): b_j of each target atom at |q|.
class ScatterFactors {
   public:
    ScatterFactors(const Database &db, std::vector<size_t> atom_ids, double background_factor = 0.0,
                   std::vector<double> kappas = {});
    const std::vector<double> &update(double ql);
    bool depends_on_q() const;
    double compute_background(double ql) const;
    size_t size() const { return ids_.size(); }

   private:
    const Database &db_;
    std::vector<size_t> ids_;
    std::vector<double> kappas_;
    double background_;
    std::vector<double> factors_;
};

// ------------------------------------------------------------ stager ------

// A source of float32 frames (NA atoms x 3), e.g. a trajectory reader.
class ICoordinateSource {
   public:
    virtual ~ICoordinateSource() = default;
    virtual size_t atoms() const = 0;
    virtual size_t frames() const = 0;
    virtual void read_frame(size_t f, float *xyz) = 0;
};

class InMemoryTrajectory : public ICoordinateSource {
   public:
    InMemoryTrajectory(size_t NA, size_t NF, std::vector<float> frames_aos);
    size_t atoms() const override { return NA_; }
    size_t frames() const override { return NF_; }
    void read_frame(size_t f, float *xyz) override;

   private:
    size_t NA_, NF_;
    std::vector<float> data_;
};

// DataStagerByFrame (data_stager.cpp:72-118): frames-major [NF][NA][3].
class DataStagerByFrame {
   public:
    DataStagerByFrame(ICoordinateSource &src, Timer &timer) : src_(src), timer_(timer) {}
    std::vector<float> stage();

   private:
    ICoordinateSource &src_;
    Timer &timer_;
};

// DataStagerByAtom (data_stager.cpp:214-338): atom-major [NA][NF][3].
class DataStagerByAtom {
   public:
    DataStagerByAtom(ICoordinateSource &src, Timer &timer) : src_(src), timer_(timer) {}
    std::vector<float> stage();

   private:
    ICoordinateSource &src_;
    Timer &timer_;
};

// ------------------------------------------------------------ output ------

// HDF5WriterClient::write (file_writer_service.hpp:64): one q shell.
class ISignalWriter {
   public:
    virtual ~ISignalWriter() = default;
    virtual void write(const CartesianCoor3D &q, const std::vector<std::complex<double>> &fqt,
                       std::complex<double> fq, std::complex<double> fq2) = 0;
    virtual void flush() {}
};

class MemorySignalWriter : public ISignalWriter {
   public:
    struct Record {
        CartesianCoor3D q;
        std::vector<std::complex<double>> fqt;
        std::complex<double> fq, fq2;
    };
    void write(const CartesianCoor3D &q, const std::vector<std::complex<double>> &fqt, std::complex<double> fq,
               std::complex<double> fq2) override;
    std::vector<Record> records;
};

// ------------------------------------------------------------ devices -----

// IScatterDevice (abstract_scatter_device.hpp:108-118)
class IScatterDevice {
   public:
    virtual ~IScatterDevice() = default;
    virtual void run() = 0;
    virtual std::map<std::string, Timer::Stat> getTimer() const = 0;
    virtual double progress() = 0;

   protected:
    virtual void runner() = 0;
    virtual size_t status() = 0;
};

// AbstractScatterDevice (abstract_scatter_device.cpp:105-244): run() stages
// once, then runner() loops compute() / write() / next() over the q shells.
// compute() evaluates a block of shells in one device call; write() hands each
// shell to the writer in q order.
class AbstractScatterDevice : public IScatterDevice {
   public:
    AbstractScatterDevice(const Params &params, ICoordinateSource &src, ScatterFactors &sf, ISignalWriter &writer,
                          std::vector<CartesianCoor3D> qvectors);
    ~AbstractScatterDevice() override;
    void run() override;
    std::map<std::string, Timer::Stat> getTimer() const override { return timer_.stats(); }
    double progress() override;

   protected:
    void runner() override;
    size_t status() override;
    virtual void stage_data() = 0;
    virtual void compute();
    void write();
    void next();
    void check(int rc, const char *what);
    // number of shells from `first` that one device call evaluates
    virtual size_t block_size(size_t first) const = 0;
    // one device call over shells [first, first + ns); bshell [ns][NA] or null
    virtual void compute_block(size_t first, size_t ns, const double *bshell, double *fqt, double *fq,
                               double *fq2) = 0;

    Params params_;
    ICoordinateSource &src_;
    ScatterFactors &scatterfactors_;
    ISignalWriter &writer_;
    std::vector<CartesianCoor3D> vectors_;
    size_t current_vector_ = 0;
    size_t NF = 0, NA = 0;
    sx_ctx *ctx_ = nullptr;
    Timer timer_;
    // results of the current block
    size_t block_first_ = 0, block_count_ = 0;
    std::vector<std::vector<std::complex<double>>> block_fqt_;
    std::vector<std::complex<double>> block_fq_, block_fq2_;
};

// The subvectors of shell q (abstract_vectors_scatter_device.cpp:112-175):
// |q| u_m for vectors of type file/sphere, the cylinder construction about
// the axis, or q itself without orientational averaging.
std::vector<CartesianCoor3D> subvectors(const ScatteringParameters &sp, const CartesianCoor3D &q);

// AbstractVectorsScatterDevice (abstract_vectors_scatter_device.cpp:31-175)
class AbstractVectorsScatterDevice : public AbstractScatterDevice {
   public:
    using AbstractScatterDevice::AbstractScatterDevice;

   protected:
    // subvectors of shell q (q_m, m < NM) as AbstractVectorsScatterDevice::init_subvectors
    std::vector<CartesianCoor3D> init_subvectors(const CartesianCoor3D &q) const;
    bool oriented() const;  // subvectors are |q| u_m for a shared set u
    size_t block_size(size_t first) const override;
    void compute_block(size_t first, size_t ns, const double *bshell, double *fqt, double *fq, double *fq2) override;
    void count_stepped(size_t ns);
};

class AllVectorsScatterDevice : public AbstractVectorsScatterDevice {
   public:
    AllVectorsScatterDevice(const Params &params, ICoordinateSource &src, ScatterFactors &sf, ISignalWriter &writer,
                            std::vector<CartesianCoor3D> qvectors);

   protected:
    void stage_data() override;
};

class SelfVectorsScatterDevice : public AbstractVectorsScatterDevice {
   public:
    SelfVectorsScatterDevice(const Params &params, ICoordinateSource &src, ScatterFactors &sf,
                             ISignalWriter &writer, std::vector<CartesianCoor3D> qvectors);

   protected:
    void stage_data() override;
};

// MPSphereScatterDevice / MPCylinderScatterDevice (multipole_scatter_device.cpp):
// the orientational average as a sum over multipole moments of the staged
// frames; q shells are batched like the vector devices.
class MultipoleScatterDevice : public AbstractScatterDevice {
   public:
    MultipoleScatterDevice(const Params &params, ICoordinateSource &src, ScatterFactors &sf, ISignalWriter &writer,
                           std::vector<CartesianCoor3D> qvectors, int kind);

   protected:
    void stage_data() override;
    size_t block_size(size_t first) const override;
    void compute_block(size_t first, size_t ns, const double *bshell, double *fqt, double *fq, double *fq2) override;
    int kind_;  // SX_MULTIPOLE_SPHERE or SX_MULTIPOLE_CYLINDER
    std::vector<int> moments_;
};

class MPSphereScatterDevice : public MultipoleScatterDevice {
   public:
    MPSphereScatterDevice(const Params &params, ICoordinateSource &src, ScatterFactors &sf, ISignalWriter &writer,
                          std::vector<CartesianCoor3D> qvectors);
};

class MPCylinderScatterDevice : public MultipoleScatterDevice {
   public:
    MPCylinderScatterDevice(const Params &params, ICoordinateSource &src, ScatterFactors &sf,
                            ISignalWriter &writer, std::vector<CartesianCoor3D> qvectors);
};

// ScatterDeviceFactory::create (scatter_device_factory.cpp:23-210)
class ScatterDeviceFactory {
   public:
    static std::unique_ptr<IScatterDevice> create(const Params &params, ICoordinateSource &src, ScatterFactors &sf,
                                                  ISignalWriter &writer, std::vector<CartesianCoor3D> qvectors);
};

}  // namespace sassena_amd
