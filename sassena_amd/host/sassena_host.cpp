// sassena_host.cpp -- host C++ layer over libsassena_amd.so (see sassena_host.hpp).
#include "sassena_host.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <random>
#include <sstream>

namespace sassena_amd {

// ------------------------------------------------------------ basics ------

double CartesianCoor3D::length() const { return std::sqrt(std::pow(x, 2) + std::pow(y, 2) + std::pow(z, 2)); }

CartesianCoor3D CartesianCoor3D::cross_product(const CartesianCoor3D &o) const {
    return {y * o.z - z * o.y, z * o.x - x * o.z, x * o.y - y * o.x};
}

CartesianCoor3D operator*(double s, const CartesianCoor3D &v) { return {s * v.x, s * v.y, s * v.z}; }
CartesianCoor3D operator/(const CartesianCoor3D &v, double s) { return {v.x / s, v.y / s, v.z / s}; }

static double now_seconds() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

void Timer::start(const std::string &label) { open_[label] = now_seconds(); }

void Timer::stop(const std::string &label) {
    auto it = open_.find(label);
    if (it == open_.end()) return;
    const double dt = now_seconds() - it->second;
    open_.erase(it);
    Stat &s = stats_[label];
    s.sum += dt;
    s.min = std::min(s.min, dt);
    s.max = std::max(s.max, dt);
    s.count++;
}

void Timer::add_count(const std::string &label, size_t n) { stats_[label].count += n; }

// ------------------------------------------------------------ control -----

void OrientationVectors::create() {
    vectors.clear();
    if (type == "file") {
        // parameters.cpp:932-944: rows "x y z", each non-zero row normalised
        std::ifstream in(filepath);
        if (!in) throw terminate_request("cannot open orientation vectors file: " + filepath);
        double x, y, z;
        while (in >> x >> y >> z) {
            CartesianCoor3D q(x, y, z);
            const double ql = q.length();
            if (ql != 0) q = (1.0 / ql) * q;
            vectors.push_back(q);
        }
    } else if (type == "sphere") {
        if (algorithm != "boost_uniform_on_sphere")
            throw terminate_request("Vectors algorithm not understood: " + algorithm);
        // Mersenne twister seeded like the reference (parameters.cpp:946-959);
        // normal triplets normalised to the sphere.  Boost's normal variate
        // algorithm is not reproduced, so the set differs from the
        // reference's for the same seed (same distribution).
        std::mt19937 rng((std::mt19937::result_type)seed);
        std::normal_distribution<double> nd(0.0, 1.0);
        for (long i = 0; i < resolution; ++i) {
            double r[3], n2 = 0;
            do {
                n2 = 0;
                for (double &c : r) {
                    c = nd(rng);
                    n2 += c * c;
                }
            } while (n2 == 0);
            const double n = std::sqrt(n2);
            vectors.emplace_back(r[0] / n, r[1] / n, r[2] / n);
        }
    } else if (type == "cylinder") {
        if (algorithm == "boost_uniform_on_sphere") {
            std::mt19937 rng((std::mt19937::result_type)seed);
            std::normal_distribution<double> nd(0.0, 1.0);
            for (long i = 0; i < resolution; ++i) {
                double a, b, n2;
                do {
                    a = nd(rng);
                    b = nd(rng);
                    n2 = a * a + b * b;
                } while (n2 == 0);
                const double n = std::sqrt(n2);
                vectors.emplace_back(a / n, b / n, 0.0);
            }
        } else if (algorithm == "raster_linear") {
            // parameters.cpp:1006-1014: phi stepped by 2 pi / (360 resolution)
            const double M_2PI = 2 * M_PI;
            const double radincr = M_2PI / (360 * resolution);
            for (double phi = 0; phi < M_2PI; phi += radincr) vectors.emplace_back(std::cos(phi), std::sin(phi), 0);
        } else {
            throw terminate_request("Vectors algorithm not understood: " + algorithm);
        }
    } else {
        throw terminate_request("Vectors orientation averaging type not understood: " + type);
    }
}

void ScatteringParameters::create_from_scans() {
    if (scans.size() > 3) throw terminate_request("More than 3 scan definitions are not supported.");
    std::vector<std::vector<CartesianCoor3D>> qs(scans.size());
    for (size_t i = 0; i < scans.size(); ++i) {
        const ScanParameters &sc = scans[i];
        if (sc.points == 0) continue;
        if (sc.points == 1) {
            qs[i].push_back(((sc.from + sc.to) / 2) * sc.basevector);
            continue;
        }
        if (sc.points == 2) {
            qs[i].push_back(sc.from * sc.basevector);
            qs[i].push_back(sc.to * sc.basevector);
            continue;
        }
        qs[i].push_back(sc.from * sc.basevector);
        for (size_t j = 1; j < sc.points - 1; ++j) {
            // single-precision powf, as parameters.cpp:1151
            const double scal =
                sc.from + powf((float)(j * 1.0 / (sc.points - 1)), (float)sc.exponent) * (sc.to - sc.from);
            qs[i].push_back(scal * sc.basevector);
        }
        qs[i].push_back(sc.to * sc.basevector);
    }
    // element-wise outer sum of up to three scans
    if (scans.size() == 1) {
        for (auto &q : qs[0]) qvectors.push_back(q);
    } else if (scans.size() == 2) {
        for (auto &a : qs[0])
            for (auto &b : qs[1]) qvectors.push_back(a + b);
    } else if (scans.size() == 3) {
        for (auto &a : qs[0])
            for (auto &b : qs[1])
                for (auto &c : qs[2]) qvectors.push_back(a + b + c);
    }
}

// ------------------------------------------------------------ database ----

void Database::add(const std::string &element, const ElementEntry &e) {
    names_.push_back(element);
    entries_.push_back(e);
}

size_t Database::id(const std::string &element) const {
    for (size_t i = 0; i < names_.size(); ++i)
        if (names_[i] == element) return i;
    throw terminate_request("element not in database: " + element);
}

static double factorial(double n) {
    double r = 1;
    for (int k = 2; k <= (int)std::lround(n); ++k) r *= k;
    return r;
}

double Database::sfactor(size_t id, double q) const {
    const DatabaseFunction &f = entries_.at(id).sfactor;
    const std::vector<double> &v = f.constants;
    switch (f.function_type) {
        case 0:  // constant (neutron scattering length), database.cpp:474-476
            return v.at(0);
        case 1: {  // Slater orbitals, database.cpp:484-511
            double den = 0.0;
            for (size_t j = 0; j < 5; ++j) {
                const size_t j2 = 5 + j, j3 = 10 + j;
                if (v.at(j) != 0.0) {
                    double c = powf(2.0 * v[j] / v[j3], v[j3] + 0.5);
                    c /= std::sqrt(factorial(2.0 * v[j3]));
                    den += v[j2] * powf(c * powf(q, v[j3] - 1.0) * std::exp(-1.0 * v[j] * q / v[j3] / 1.0), 2);
                }
            }
            return den / (4.0 * M_PI);
        }
        case 2: {  // four Gaussians plus constant, database.cpp:514-523
            double den = 0.0;
            den += v.at(0) * std::exp(-v.at(1) * q * q);
            den += v.at(2) * std::exp(-v.at(3) * q * q);
            den += v.at(4) * std::exp(-v.at(5) * q * q);
            den += v.at(6) * std::exp(-v.at(7) * q * q);
            den += v.at(8);
            return den;
        }
        default:
            throw terminate_request("ScatterFactor-type not implemented: type=" + std::to_string(f.function_type));
    }
}

double Database::volume(size_t id) const {
    const DatabaseFunction &f = entries_.at(id).volume;
    const std::vector<double> &v = f.constants;
    switch (f.function_type) {  // database.cpp:391-410
        case 0: return v.empty() ? 0.0 : v[0];
        case 1: return (4.0 / 3.0) * M_PI * powf(v.at(0), 3);
        case 2: return std::sqrt(powf(M_PI, 3)) * powf(v.at(0), 3);
        default: throw terminate_request("Size-type not implemented: type=" + std::to_string(f.function_type));
    }
}

double Database::exclusion(size_t id, double effvolume, double q) const {
    const DatabaseFunction &f = entries_.at(id).exclusion;
    const std::vector<double> &v = f.constants;
    switch (f.function_type) {  // database.cpp:430-446
        case 0: return v.empty() ? 0.0 : v[0];
        case 1: return effvolume * v.at(0);
        case 2: return effvolume * std::exp(-1.0 * powf(effvolume, 2.0 / 3.0) * powf(q, 2) / (4 * M_PI)) * v.at(0);
        default:
            throw terminate_request("ExclusionParameter-type not implemented: type=" + std::to_string(f.function_type));
    }
}

Database Database::neutron_coherent() {
    // bound coherent scattering lengths in fm (Sears, Neutron News 3 (1992) 26)
    static const std::pair<const char *, double> table[] = {
        {"H", -3.7390}, {"D", 6.671}, {"C", 6.6460}, {"N", 9.36}, {"O", 5.803}, {"P", 5.13}, {"S", 2.847},
    };
    Database db;
    for (auto &e : table) {
        ElementEntry en;
        en.sfactor.function_type = 0;
        en.sfactor.constants = {e.second};
        en.volume.constants = {0.0};
        en.exclusion.constants = {0.0};
        db.add(e.first, en);
    }
    return db;
}

namespace {
// shortest decimal form that reads back as the same double
std::string shortest(double v) {
    char buf[40];
    for (int p = 6; p <= 17; ++p) {
        std::snprintf(buf, sizeof buf, "%.*g", p, v);
        if (std::strtod(buf, nullptr) == v) break;
    }
    return buf;
}
}  // namespace

std::string Database::to_xml() const {
    std::ostringstream o;
    o << "<?xml version=\"1.0\"?>\n<root>\n";
    const char *groups[3] = {"scatterfactors", "sizes", "exclusionfactors"};
    for (int g = 0; g < 3; ++g) {
        o << "  <" << groups[g] << ">\n";
        for (size_t k = 0; k < names_.size(); ++k) {
            const DatabaseFunction &f = g == 0 ? entries_[k].sfactor : g == 1 ? entries_[k].volume
                                                                              : entries_[k].exclusion;
            o << "    <element>\n      <name>" << names_[k] << "</name>\n      <type>" << f.function_type
              << "</type>\n";
            for (double c : f.constants) o << "      <param>" << shortest(c) << "</param>\n";
            o << "    </element>\n";
        }
        o << "  </" << groups[g] << ">\n";
    }
    o << "</root>\n";
    return o.str();
}

ScatterFactors::ScatterFactors(const Database &db, std::vector<size_t> atom_ids, double background_factor,
                               std::vector<double> kappas)
    : db_(db), ids_(std::move(atom_ids)), kappas_(std::move(kappas)), background_(background_factor) {
    if (kappas_.empty()) kappas_.assign(ids_.size(), 1.0);  // scatter_factors.cpp:92
    if (kappas_.size() != ids_.size()) throw terminate_request("ScatterFactors: kappas do not match atoms");
    factors_.resize(ids_.size());
}

const std::vector<double> &ScatterFactors::update(double ql) {
    // scatter_factors.cpp:56-78: sf - background * exclusion(kappa * volume, |q|)
    for (size_t i = 0; i < ids_.size(); ++i) {
        double sf = db_.sfactor(ids_[i], ql);
        const double v = db_.volume(ids_[i]);
        const double efactor = db_.exclusion(ids_[i], kappas_[i] * v, ql);
        sf = sf - background_ * efactor;
        factors_[i] = sf;
    }
    return factors_;
}

bool ScatterFactors::depends_on_q() const {
    std::vector<double> a, b;
    ScatterFactors *self = const_cast<ScatterFactors *>(this);
    const std::vector<double> keep = factors_;
    a = self->update(0.37);
    b = self->update(2.91);
    self->factors_ = keep;
    return a != b;
}

double ScatterFactors::compute_background(double ql) const {
    // scatter_factors.cpp:104-123
    double efactor_sum = 0, sf_sum = 0;
    for (size_t i = 0; i < ids_.size(); ++i) {
        const double sf = db_.sfactor(ids_[i], ql);
        const double v = db_.volume(ids_[i]);
        efactor_sum += db_.exclusion(ids_[i], kappas_[i] * v, ql);
        sf_sum += sf;
    }
    return sf_sum / efactor_sum;
}

// ------------------------------------------------------------ stager ------

InMemoryTrajectory::InMemoryTrajectory(size_t NA, size_t NF, std::vector<float> frames_aos)
    : NA_(NA), NF_(NF), data_(std::move(frames_aos)) {
    if (data_.size() != NA * NF * 3) throw terminate_request("InMemoryTrajectory: size mismatch");
}

void InMemoryTrajectory::read_frame(size_t f, float *xyz) {
    std::copy(data_.begin() + f * NA_ * 3, data_.begin() + (f + 1) * NA_ * 3, xyz);
}

std::vector<float> DataStagerByFrame::stage() {
    const size_t NA = src_.atoms(), NF = src_.frames();
    timer_.start("st:load");
    std::vector<float> out(NA * NF * 3);
    for (size_t f = 0; f < NF; ++f) src_.read_frame(f, out.data() + f * NA * 3);
    timer_.stop("st:load");
    return out;
}

std::vector<float> DataStagerByAtom::stage() {
    const size_t NA = src_.atoms(), NF = src_.frames();
    timer_.start("st:load");
    std::vector<float> frame(NA * 3), out(NA * NF * 3);
    for (size_t f = 0; f < NF; ++f) {
        src_.read_frame(f, frame.data());
        // transpose into p_coordinates[(a*NF + f)*3 + c] (data_stager.cpp:327-330)
        for (size_t a = 0; a < NA; ++a)
            for (int c = 0; c < 3; ++c) out[(a * NF + f) * 3 + c] = frame[a * 3 + c];
    }
    timer_.stop("st:load");
    return out;
}

void MemorySignalWriter::write(const CartesianCoor3D &q, const std::vector<std::complex<double>> &fqt,
                               std::complex<double> fq, std::complex<double> fq2) {
    records.push_back({q, fqt, fq, fq2});
}

// ------------------------------------------------------------ devices -----

static int dsp_code(const std::string &s) {
    if (s == "autocorrelate") return SX_DSP_AUTOCORRELATE;
    if (s == "square") return SX_DSP_SQUARE;
    if (s == "plain") return SX_DSP_PLAIN;
    throw terminate_request("DSP type not understood: " + s);
}

static int method_code(const std::string &s) {
    if (s == "fftw") return SX_METHOD_FFT;
    if (s == "direct") return SX_METHOD_DIRECT;
    throw terminate_request("Correlation method not understood: " + s);
}

AbstractScatterDevice::AbstractScatterDevice(const Params &params, ICoordinateSource &src, ScatterFactors &sf,
                                             ISignalWriter &writer, std::vector<CartesianCoor3D> qvectors)
    : params_(params), src_(src), scatterfactors_(sf), writer_(writer), vectors_(std::move(qvectors)) {
    NF = src.frames();
    NA = src.atoms();
    if (NF < 1) throw terminate_request("No frames available. Aborting");
    if (NA < 1) throw terminate_request("No atoms available. Aborting");
    if (vectors_.empty()) throw terminate_request("No qvectors left to compute. Aborting");
    if (sf.size() != NA) throw terminate_request("scattering factors do not match the target atoms");
    sx_config cfg{};
    cfg.device = params.device;
    cfg.scattering_type = params.scattering.type == "self" ? SX_SCATTERING_SELF : SX_SCATTERING_ALL;
    cfg.dsp_type = dsp_code(params.scattering.dsp_type);
    cfg.dsp_method = method_code(params.scattering.dsp_method);
    const int rc = sx_create(&cfg, &ctx_);
    if (rc != SX_OK) throw device_error(rc, "sx_create failed (" + std::to_string(rc) + ")");
}

AbstractScatterDevice::~AbstractScatterDevice() { sx_destroy(ctx_); }

void AbstractScatterDevice::check(int rc, const char *what) {
    if (rc == SX_OK) return;
    // device memory exhausted ends the run as the reference's memory checks
    // do (abstract_scatter_device.cpp:111)
    if (rc == SX_ERR_NOMEM) throw terminate_request(std::string(what) + ": " + sx_last_error(ctx_));
    throw device_error(rc, std::string(what) + ": " + sx_last_error(ctx_));
}

void AbstractScatterDevice::run() {
    timer_.start("sd:stage");
    stage_data();
    timer_.stop("sd:stage");
    timer_.start("sd:runner");
    runner();
    timer_.stop("sd:runner");
    writer_.flush();
}

void AbstractScatterDevice::runner() {
    while (status() == 0) {
        timer_.start("sd:compute");
        compute();
        timer_.stop("sd:compute");
        timer_.start("sd:write");
        write();
        timer_.stop("sd:write");
        next();
    }
}

size_t AbstractScatterDevice::status() { return current_vector_ == vectors_.size() ? 1 : 0; }

double AbstractScatterDevice::progress() { return (double)current_vector_ / (double)vectors_.size(); }

void AbstractScatterDevice::next() { ++current_vector_; }

void AbstractScatterDevice::write() {
    const size_t k = current_vector_ - block_first_;
    writer_.write(vectors_[current_vector_], block_fqt_[k], block_fq_[k], block_fq2_[k]);
}

// Evaluates the block of shells starting at current_vector_ when its results
// are not cached: consecutive shells with the same number of subvectors go
// into one device call (up to params.batch_shells).
void AbstractScatterDevice::compute() {
    if (current_vector_ >= block_first_ && current_vector_ < block_first_ + block_count_) return;
    const size_t first = current_vector_;
    const size_t ns = block_size(first);
    // scattering factors: one set when they do not depend on |q|, else per shell
    std::vector<double> bshell;
    const bool per_shell = scatterfactors_.depends_on_q();
    if (per_shell) {
        bshell.resize(ns * NA);
        for (size_t s = 0; s < ns; ++s) {
            const std::vector<double> &b = scatterfactors_.update(vectors_[first + s].length());
            std::copy(b.begin(), b.end(), bshell.begin() + s * NA);
        }
    } else {
        const std::vector<double> &b = scatterfactors_.update(vectors_[first].length());
        check(sx_set_factors(ctx_, b.data(), NA), "sx_set_factors");
    }
    std::vector<double> fqt(ns * NF * 2), fq(ns * 2), fq2(ns * 2);
    compute_block(first, ns, per_shell ? bshell.data() : nullptr, fqt.data(), fq.data(), fq2.data());
    block_first_ = first;
    block_count_ = ns;
    block_fqt_.assign(ns, std::vector<std::complex<double>>(NF));
    block_fq_.resize(ns);
    block_fq2_.resize(ns);
    for (size_t s = 0; s < ns; ++s) {
        for (size_t t = 0; t < NF; ++t) block_fqt_[s][t] = {fqt[(s * NF + t) * 2], fqt[(s * NF + t) * 2 + 1]};
        block_fq_[s] = {fq[2 * s], fq[2 * s + 1]};
        block_fq2_[s] = {fq2[2 * s], fq2[2 * s + 1]};
    }
}

// abstract_vectors_scatter_device.cpp:112-175
std::vector<CartesianCoor3D> subvectors(const ScatteringParameters &sp, const CartesianCoor3D &q) {
    std::vector<CartesianCoor3D> out;
    const auto &vecs = sp.vectors.vectors;
    if (sp.orientation_type == "vectors" && !vecs.empty()) {
        if (sp.vectors.type == "file" || sp.vectors.type == "sphere") {
            const double ql = q.length();
            for (const auto &u : vecs) out.push_back(ql * u);
        } else if (sp.vectors.type == "cylinder") {
            // base from the axis (coor3d.cpp:278-296), q projected on it
            const CartesianCoor3D ek(0, 0, 1), ej(0, 1, 0);
            const CartesianCoor3D ez = sp.axis / sp.axis.length();
            CartesianCoor3D ekez = ek.cross_product(ez);
            if (ekez.length() == 0) ekez = ej.cross_product(ez);
            const CartesianCoor3D er = ekez / ekez.length();
            const CartesianCoor3D ezer = ez.cross_product(er);
            const CartesianCoor3D ephi = ezer / ezer.length();
            const CartesianCoor3D qp(q.dot(er), q.dot(ephi), q.dot(ez));
            const double r = std::sqrt(std::pow(qp.x, 2) + std::pow(qp.y, 2));
            if (r == 0) {
                out.push_back(qp);
            } else {
                for (const auto &v : vecs) out.push_back(qp.z * ez + r * (v.x * er + v.y * ephi));
            }
        } else {
            throw terminate_request("Vectors orientation averaging type not understood: " + sp.vectors.type);
        }
    } else {
        out.push_back(q);
    }
    return out;
}

std::vector<CartesianCoor3D> AbstractVectorsScatterDevice::init_subvectors(const CartesianCoor3D &q) const {
    return subvectors(params_.scattering, q);
}

bool AbstractVectorsScatterDevice::oriented() const {
    const ScatteringParameters &sp = params_.scattering;
    return sp.orientation_type == "vectors" && !sp.vectors.vectors.empty() &&
           (sp.vectors.type == "file" || sp.vectors.type == "sphere");
}

AllVectorsScatterDevice::AllVectorsScatterDevice(const Params &params, ICoordinateSource &src, ScatterFactors &sf,
                                                 ISignalWriter &writer, std::vector<CartesianCoor3D> qvectors)
    : AbstractVectorsScatterDevice(params, src, sf, writer, std::move(qvectors)) {}

void AllVectorsScatterDevice::stage_data() {
    // all_vectors_scatter_device.cpp:127-132: frames-major staging
    DataStagerByFrame stager(src_, timer_);
    std::vector<float> block = stager.stage();
    timer_.start("st:device");
    check(sx_stage(ctx_, block.data(), NA, NF, SX_LAYOUT_FRAMES_AOS), "sx_stage");
    timer_.stop("st:device");
}

SelfVectorsScatterDevice::SelfVectorsScatterDevice(const Params &params, ICoordinateSource &src, ScatterFactors &sf,
                                                   ISignalWriter &writer, std::vector<CartesianCoor3D> qvectors)
    : AbstractVectorsScatterDevice(params, src, sf, writer, std::move(qvectors)) {}

void SelfVectorsScatterDevice::stage_data() {
    // self path stages atom-major (data_stager.cpp:214-338)
    DataStagerByAtom stager(src_, timer_);
    std::vector<float> block = stager.stage();
    timer_.start("st:device");
    check(sx_stage(ctx_, block.data(), NA, NF, SX_LAYOUT_ATOMS_AOS), "sx_stage");
    timer_.stop("st:device");
}

// Shells of the last oriented call that stepped phasors from shell to shell
// ("sd:stepped"; "sd:stepped-residual" when corrected for the single-precision
// rounding of a q scan): counts only, for the run report.
void AbstractVectorsScatterDevice::count_stepped(size_t ns) {
    sx_timing tm{};
    if (sx_get_timing(ctx_, &tm) != SX_OK || !tm.shell_progression) return;
    timer_.add_count(tm.shell_residuals ? "sd:stepped-residual" : "sd:stepped", ns);
}

size_t AbstractVectorsScatterDevice::block_size(size_t first) const {
    // consecutive shells with the same number of subvectors share one call
    const size_t NM = init_subvectors(vectors_[first]).size();
    size_t ns = 1;
    while (ns < params_.batch_shells && first + ns < vectors_.size() &&
           init_subvectors(vectors_[first + ns]).size() == NM)
        ++ns;
    return ns;
}

void AbstractVectorsScatterDevice::compute_block(size_t first, size_t ns, const double *bshell, double *fqt,
                                                 double *fq, double *fq2) {
    if (oriented()) {
        // q_{s,m} = |q_s| u_m with the shared orientation set
        const auto &u = params_.scattering.vectors.vectors;
        std::vector<double> uu(3 * u.size()), ql(ns);
        for (size_t m = 0; m < u.size(); ++m) {
            uu[3 * m] = u[m].x;
            uu[3 * m + 1] = u[m].y;
            uu[3 * m + 2] = u[m].z;
        }
        for (size_t s = 0; s < ns; ++s) ql[s] = vectors_[first + s].length();
        check(sx_compute_shells_oriented(ctx_, ns, ql.data(), uu.data(), u.size(), bshell, fqt, fq, fq2),
              "sx_compute_shells_oriented");
        count_stepped(ns);
        return;
    }
    const size_t NM = init_subvectors(vectors_[first]).size();
    std::vector<double> qv(ns * NM * 3);
    for (size_t s = 0; s < ns; ++s) {
        const std::vector<CartesianCoor3D> sub = init_subvectors(vectors_[first + s]);
        for (size_t m = 0; m < NM; ++m) {
            qv[(s * NM + m) * 3] = sub[m].x;
            qv[(s * NM + m) * 3 + 1] = sub[m].y;
            qv[(s * NM + m) * 3 + 2] = sub[m].z;
        }
    }
    // Shells whose subvectors are |q_s| times one shared set (a scan along a
    // fixed direction: orientation none, or cylinder averaging about a fixed
    // axis) go through the oriented entry point as well.
    std::vector<double> ql(ns), uu(3 * NM);
    bool shared = true;
    for (size_t s = 0; s < ns && shared; ++s) {
        ql[s] = vectors_[first + s].length();
        shared = ql[s] > 0.0;
    }
    for (size_t m = 0; m < NM && shared; ++m)
        for (int k = 0; k < 3; ++k) uu[3 * m + k] = qv[3 * m + k] / ql[0];
    for (size_t s = 1; s < ns && shared; ++s)
        for (size_t m = 0; m < NM && shared; ++m)
            for (int k = 0; k < 3 && shared; ++k) {
                const double want = ql[s] * uu[3 * m + k], got = qv[(s * NM + m) * 3 + k];
                shared = std::fabs(got - want) <= 1e-12 * ql[s];
            }
    if (shared) {
        check(sx_compute_shells_oriented(ctx_, ns, ql.data(), uu.data(), NM, bshell, fqt, fq, fq2),
              "sx_compute_shells_oriented");
        count_stepped(ns);
        return;
    }
    check(sx_compute_shells(ctx_, ns, qv.data(), NM, bshell, fqt, fq, fq2), "sx_compute_shells");
}

void MultipoleMoments::create(const std::string &multipole_type) {
    moments.clear();
    if (type == "file") {
        std::ifstream in(filepath);
        if (!in) throw terminate_request("cannot open multipole moments file: " + filepath);
        long major, minor;
        while (in >> major >> minor) moments.emplace_back(major, minor);
    } else if (type == "resolution") {
        if (multipole_type == "sphere") {
            moments.emplace_back(0L, 0L);
            for (long l = 1; l <= resolution; ++l)
                for (long m = -l; m <= l; ++m) moments.emplace_back(l, m);
        } else if (multipole_type == "cylinder") {
            moments.emplace_back(0L, 0L);
            for (long l = 1; l <= resolution; ++l)
                for (long m = 0; m < 4; ++m) moments.emplace_back(l, m);
        } else {
            throw terminate_request("Type not understood: scattering.average.orientation.multipole.type=" +
                                    multipole_type);
        }
    } else {
        throw terminate_request("Type not understood: scattering.average.orientation.multipole.moments.type=" + type);
    }
    for (const auto &mm : moments) {
        const std::string id = "major=" + std::to_string(mm.first) + ", minor=" + std::to_string(mm.second);
        if (mm.first < 0) throw terminate_request("Major multipole moment must be >= 0! " + id);
        if (multipole_type == "cylinder") {
            if (mm.second < 0 || mm.second > 3)
                throw terminate_request("Minor multipole moment must be between 0 and 3! " + id);
            if (mm.first == 0 && mm.second != 0)
                throw terminate_request("Minor multipole moment must be 0 for Major 0! " + id);
        } else if (multipole_type == "sphere") {
            if (std::labs(mm.second) > mm.first)
                throw terminate_request("Minor multipole moment must be between -Major and +Major! " + id);
        }
    }
}

MultipoleScatterDevice::MultipoleScatterDevice(const Params &params, ICoordinateSource &src, ScatterFactors &sf,
                                               ISignalWriter &writer, std::vector<CartesianCoor3D> qvectors,
                                               int kind)
    : AbstractScatterDevice(params, src, sf, writer, std::move(qvectors)), kind_(kind) {
    const auto &mm = params_.scattering.moments.moments;
    if (mm.empty()) throw terminate_request("multipole orientational averaging needs at least one moment");
    for (const auto &p : mm) {
        moments_.push_back((int)p.first);
        moments_.push_back((int)p.second);
    }
}

void MultipoleScatterDevice::stage_data() {
    DataStagerByFrame stager(src_, timer_);
    std::vector<float> block = stager.stage();
    timer_.start("st:device");
    check(sx_stage(ctx_, block.data(), NA, NF, SX_LAYOUT_FRAMES_AOS), "sx_stage");
    timer_.stop("st:device");
}

size_t MultipoleScatterDevice::block_size(size_t first) const {
    return std::min(params_.batch_shells, vectors_.size() - first);
}

void MultipoleScatterDevice::compute_block(size_t first, size_t ns, const double *bshell, double *fqt, double *fq,
                                           double *fq2) {
    std::vector<double> qv(3 * ns);
    for (size_t s = 0; s < ns; ++s) {
        qv[3 * s] = vectors_[first + s].x;
        qv[3 * s + 1] = vectors_[first + s].y;
        qv[3 * s + 2] = vectors_[first + s].z;
    }
    const CartesianCoor3D &ax = params_.scattering.axis;
    const double axis[3] = {ax.x, ax.y, ax.z};
    check(sx_compute_multipole(ctx_, kind_, ns, qv.data(), axis, moments_.data(), moments_.size() / 2, bshell, fqt,
                               fq, fq2),
          "sx_compute_multipole");
}

MPSphereScatterDevice::MPSphereScatterDevice(const Params &params, ICoordinateSource &src, ScatterFactors &sf,
                                             ISignalWriter &writer, std::vector<CartesianCoor3D> qvectors)
    : MultipoleScatterDevice(params, src, sf, writer, std::move(qvectors), SX_MULTIPOLE_SPHERE) {}

MPCylinderScatterDevice::MPCylinderScatterDevice(const Params &params, ICoordinateSource &src, ScatterFactors &sf,
                                                 ISignalWriter &writer, std::vector<CartesianCoor3D> qvectors)
    : MultipoleScatterDevice(params, src, sf, writer, std::move(qvectors), SX_MULTIPOLE_CYLINDER) {}

void ram_check(const Params &params, size_t NA, size_t NF) {
    const MemoryLimits &m = params.limits;
    const size_t result = NF * 2 * sizeof(double);
    if (result > m.scale * m.result_buffer)
        throw terminate_request("limits.computation.memory.result_buffer too small: limits.computation.memory."
                                "result_buffer=" + std::to_string(m.result_buffer) + ", limits.computation.memory."
                                "scale=" + std::to_string(m.scale) + ", requested: " + std::to_string(result));
    const size_t data = NA * NF * 3 * sizeof(float);
    if (data > m.stage_data)
        throw terminate_request("Insufficient Buffer size for coordinates (limits.stage.memory.data=" +
                                std::to_string(m.stage_data) + "), requested (bytes): " + std::to_string(data));
}

std::unique_ptr<IScatterDevice> ScatterDeviceFactory::create(const Params &params, ICoordinateSource &src,
                                                             ScatterFactors &sf, ISignalWriter &writer,
                                                             std::vector<CartesianCoor3D> qvectors) {
    const ScatteringParameters &sp = params.scattering;
    // memory limits before the device is created and anything is staged
    ram_check(params, src.atoms(), src.frames());
    if (sp.type == "self") return std::make_unique<SelfVectorsScatterDevice>(params, src, sf, writer, qvectors);
    if (sp.type == "all") {
        if (sp.orientation_type == "vectors" || sp.orientation_type == "none")
            return std::make_unique<AllVectorsScatterDevice>(params, src, sf, writer, qvectors);
        if (sp.orientation_type == "multipole") {
            if (sp.multipole_type == "sphere")
                return std::make_unique<MPSphereScatterDevice>(params, src, sf, writer, qvectors);
            if (sp.multipole_type == "cylinder")
                return std::make_unique<MPCylinderScatterDevice>(params, src, sf, writer, qvectors);
            throw terminate_request("scattering.average.orientation.multipole.type not understood: " +
                                    sp.multipole_type);
        }
        throw terminate_request("Orientation averaging type not understood: " + sp.orientation_type);
    }
    throw terminate_request("scattering.type not understood: " + sp.type);
}

}  // namespace sassena_amd
