// sassena_io.cpp -- structure, trajectory and signal-file I/O of the hot path.
#include "sassena_io.hpp"
#include "sassena_config.hpp"

#include <hdf5.h>
#include <hdf5_hl.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <set>
#include <sstream>
#include <tuple>

namespace sassena_amd {

// ------------------------------------------------------------ structure ---

std::vector<std::string> read_structure(const std::string &file, const std::string &format) {
    std::ifstream in(file);
    if (!in) throw terminate_request("cannot open structure file: " + file);
    std::vector<std::string> out;
    std::string line;
    if (format == "elements") {
        while (std::getline(in, line)) {
            std::stringstream ss(line);
            std::string e;
            if (ss >> e && e[0] != '#') out.push_back(e);
        }
    } else if (format == "pdb") {
        while (std::getline(in, line)) {
            if (line.compare(0, 4, "ATOM") != 0 && line.compare(0, 6, "HETATM") != 0) continue;
            std::string e;
            if (line.size() >= 78) {
                e = line.substr(76, 2);
                e.erase(std::remove(e.begin(), e.end(), ' '), e.end());
            }
            if (e.empty() && line.size() >= 16) {
                std::string name = line.substr(12, 4);
                name.erase(std::remove(name.begin(), name.end(), ' '), name.end());
                size_t k = 0;
                while (k < name.size() && std::isdigit((unsigned char)name[k])) ++k;
                if (k < name.size()) e = name.substr(k, 1);
            }
            if (e.empty()) throw terminate_request("pdb atom without element: " + line);
            if (e.size() == 2) e[1] = (char)std::tolower((unsigned char)e[1]);
            out.push_back(e);
        }
    } else {
        throw terminate_request("structure format not understood: " + format);
    }
    if (out.empty()) throw terminate_request("no atoms in structure file: " + file);
    return out;
}

// ------------------------------------------------------------ DCD ---------

bool DCDTrajectory::detect(const std::string &file) {
    // frames.cpp:372-383: record marker 84, "CORD", closing marker at byte 88
    std::ifstream in(file, std::ios::binary);
    char buf[92];
    if (!in.read(buf, 92)) return false;
    const char m84[4] = {0x54, 0, 0, 0}, cord[4] = {'C', 'O', 'R', 'D'};
    return !std::memcmp(buf, m84, 4) && !std::memcmp(buf + 88, m84, 4) && !std::memcmp(buf + 4, cord, 4);
}

DCDTrajectory::DCDTrajectory(const std::string &file) : file_(file) {
    if (!detect(file)) throw terminate_request("file '" + file + "' appears to not be a DCD file");
    std::ifstream in(file, std::ios::binary);
    int32_t head[23];
    in.read(reinterpret_cast<char *>(head), sizeof head);
    const int32_t nframes = head[3];     // NSET
    const int32_t has_cell = head[12];   // unit-cell block flag
    const int32_t has_4d = head[13];     // fourth-dimension block flag
    int32_t marker = 0;
    in.seekg(23 * 4, std::ios::beg);
    in.read(reinterpret_cast<char *>(&marker), 4);  // title record
    in.seekg(marker, std::ios::cur);
    in.read(reinterpret_cast<char *>(&marker), 4);
    uint32_t noa = 0;
    in.read(reinterpret_cast<char *>(&marker), 4);
    in.read(reinterpret_cast<char *>(&noa), 4);
    in.read(reinterpret_cast<char *>(&marker), 4);
    if (!in) throw terminate_request("truncated DCD header: " + file);
    NA_ = noa;
    first_ = (long long)in.tellg();
    long long pos = 0;
    auto skip_record = [&](long long &p) {
        int32_t m = 0;
        in.seekg(first_ + p, std::ios::beg);
        in.read(reinterpret_cast<char *>(&m), 4);
        if (!in) throw terminate_request("truncated DCD frame: " + file);
        const long long data = p + 4;
        p += 8 + m;
        return data;
    };
    if (has_cell) skip_record(pos);
    x_off_ = skip_record(pos);
    y_off_ = skip_record(pos);
    z_off_ = skip_record(pos);
    if (has_4d) skip_record(pos);
    block_ = pos;
    // frame count from the file size (the header's NSET may lag behind)
    in.clear();
    in.seekg(0, std::ios::end);
    const long long size = (long long)in.tellg();
    const long long nfile = (size - first_) / block_;
    NF_ = (size_t)(nframes > 0 ? std::min<long long>(nframes, nfile) : nfile);
    buf_.resize(NA_);
}

void DCDTrajectory::read_frame(size_t f, float *xyz) {
    if (f >= NF_) throw terminate_request("DCD frame out of range");
    std::ifstream in(file_, std::ios::binary);
    const long long base = first_ + (long long)f * block_;
    const long long offs[3] = {x_off_, y_off_, z_off_};
    for (int c = 0; c < 3; ++c) {
        in.seekg(base + offs[c], std::ios::beg);
        in.read(reinterpret_cast<char *>(buf_.data()), sizeof(float) * NA_);
        if (!in) throw terminate_request("truncated DCD frame: " + file_);
        for (size_t a = 0; a < NA_; ++a) xyz[3 * a + c] = buf_[a];
    }
}

void write_dcd(const std::string &file, size_t NA, size_t NF, const float *frames) {
    std::ofstream out(file, std::ios::binary);
    auto rec = [&](const void *p, int32_t n) {
        out.write(reinterpret_cast<const char *>(&n), 4);
        out.write(reinterpret_cast<const char *>(p), n);
        out.write(reinterpret_cast<const char *>(&n), 4);
    };
    int32_t hdr[21] = {0};
    std::memcpy(&hdr[0], "CORD", 4);
    hdr[1] = (int32_t)NF;  // NSET
    hdr[2] = 0;            // ISTART
    hdr[3] = 1;            // NSAVC
    hdr[20] = 24;          // CHARMM version
    rec(hdr, 84);
    char title[4 + 80] = {0};
    int32_t ntitle = 1;
    std::memcpy(title, &ntitle, 4);
    std::memset(title + 4, ' ', 80);
    std::memcpy(title + 4, "sassena_amd test trajectory", 27);
    rec(title, 84);
    int32_t na = (int32_t)NA;
    rec(&na, 4);
    std::vector<float> c(NA);
    for (size_t f = 0; f < NF; ++f)
        for (int k = 0; k < 3; ++k) {
            for (size_t a = 0; a < NA; ++a) c[a] = frames[(f * NA + a) * 3 + k];
            rec(c.data(), (int32_t)(sizeof(float) * NA));
        }
    if (!out) throw terminate_request("cannot write DCD file: " + file);
}

// ------------------------------------------------------------ raw ---------

RawTrajectory::RawTrajectory(const std::string &file) : file_(file) {
    std::ifstream in(file, std::ios::binary);
    char magic[6];
    uint64_t dims[2];
    if (!in.read(magic, 6) || std::memcmp(magic, "SXRAW1", 6) || !in.read(reinterpret_cast<char *>(dims), 16))
        throw terminate_request("not a raw trajectory: " + file);
    NA_ = dims[0];
    NF_ = dims[1];
}

void RawTrajectory::read_frame(size_t f, float *xyz) {
    std::ifstream in(file_, std::ios::binary);
    in.seekg(22 + (long long)(f * NA_ * 3 * sizeof(float)), std::ios::beg);
    if (!in.read(reinterpret_cast<char *>(xyz), sizeof(float) * 3 * NA_))
        throw terminate_request("truncated raw trajectory: " + file_);
}

void RawTrajectory::write(const std::string &file, size_t NA, size_t NF, const float *frames) {
    std::ofstream out(file, std::ios::binary);
    uint64_t dims[2] = {NA, NF};
    out.write("SXRAW1", 6);
    out.write(reinterpret_cast<const char *>(dims), 16);
    out.write(reinterpret_cast<const char *>(frames), sizeof(float) * 3 * NA * NF);
}

std::unique_ptr<ICoordinateSource> open_trajectory(const std::string &file, const std::string &format) {
    if (format == "dcd") return std::make_unique<DCDTrajectory>(file);
    if (format == "raw") return std::make_unique<RawTrajectory>(file);
    throw terminate_request("trajectory format not understood: " + format);
}

// ------------------------------------------------------------ HDF5 --------

namespace {

struct H5Quiet {
    H5E_auto2_t fn = nullptr;
    void *data = nullptr;
    H5Quiet() {
        H5Eget_auto2(H5E_DEFAULT, &fn, &data);
        H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    }
    ~H5Quiet() { H5Eset_auto2(H5E_DEFAULT, fn, data); }
};

void h5check(herr_t rc, const char *what) {
    if (rc < 0) throw terminate_request(std::string("HDF5 error: ") + what);
}

void create_extendible(hid_t file, const char *name, int rank, const hsize_t *inner, hsize_t chunk0) {
    hsize_t dims[3] = {0, 0, 0}, mdims[3] = {H5S_UNLIMITED, 0, 0}, cdims[3] = {chunk0, 0, 0};
    for (int k = 1; k < rank; ++k) {
        dims[k] = mdims[k] = inner[k - 1];
        cdims[k] = inner[k - 1];
    }
    const double fill = 0;
    hid_t dcpl = H5Pcreate(H5P_DATASET_CREATE);
    h5check(H5Pset_chunk(dcpl, rank, cdims), "set_chunk");
    H5Pset_fill_value(dcpl, H5T_NATIVE_DOUBLE, &fill);
    hid_t space = H5Screate_simple(rank, dims, mdims);
    hid_t ds = H5Dcreate2(file, name, H5T_NATIVE_DOUBLE, space, H5P_DEFAULT, dcpl, H5P_DEFAULT);
    H5Sclose(space);
    H5Pclose(dcpl);
    if (ds < 0) throw terminate_request(std::string("cannot create dataset ") + name);
    H5Dclose(ds);
}

void append_rows(hid_t file, const char *name, int rank, const std::vector<double> &data, size_t nq) {
    if (nq == 0) return;
    hid_t ds = H5Dopen2(file, name, H5P_DEFAULT);
    if (ds < 0) return;  // dataset disabled in this file
    hid_t space = H5Dget_space(ds);
    hsize_t dims[3] = {0, 0, 0};
    H5Sget_simple_extent_dims(space, dims, nullptr);
    H5Sclose(space);
    const hsize_t old = dims[0];
    dims[0] += nq;
    h5check(H5Dset_extent(ds, dims), "set_extent");
    space = H5Dget_space(ds);
    hsize_t start[3] = {old, 0, 0}, count[3] = {nq, dims[1], dims[2]};
    h5check(H5Sselect_hyperslab(space, H5S_SELECT_SET, start, nullptr, count, nullptr), "select");
    hid_t mspace = H5Screate_simple(rank, count, nullptr);
    h5check(H5Dwrite(ds, H5T_NATIVE_DOUBLE, mspace, space, H5P_DEFAULT, data.data()), "write");
    H5Sclose(mspace);
    H5Sclose(space);
    H5Dclose(ds);
}

}  // namespace

HDF5SignalWriter::HDF5SignalWriter(const std::string &file, size_t NF, const RunFiles &flags,
                                   const std::string &rawconfig, const std::string &config_doc,
                                   const std::string &database_doc)
    : file_(file), NF_(NF), flags_(flags) {
    H5Quiet quiet;
    hid_t h = H5Fopen(file.c_str(), H5F_ACC_RDWR, H5P_DEFAULT);
    if (h >= 0) {
        // existing file: its fqt must have NF frames to be appended to
        if (H5LTfind_dataset(h, "fqt") == 1) {
            hsize_t dims[3];
            H5LTget_dataset_info(h, "fqt", dims, nullptr, nullptr);
            if (dims[1] != NF) {
                H5Fclose(h);
                throw terminate_request("signal file " + file + " has fqt with " + std::to_string(dims[1]) +
                                        " frames, this run has " + std::to_string(NF));
            }
        }
        H5Fclose(h);
        return;
    }
    h = H5Fcreate(file.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (h < 0) throw terminate_request("cannot create signal file: " + file);
    hid_t meta = H5Gcreate2(h, "meta", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    hsize_t rawsize = rawconfig.size();
    if (rawsize) H5LTmake_dataset_char(meta, "rawconfig", 1, &rawsize, rawconfig.data());
    // string datasets, as the reference writes them (file_writer_service.cpp:58-62)
    H5LTmake_dataset_string(meta, "config", config_doc.c_str());
    H5LTmake_dataset_string(meta, "database", database_doc.c_str());
    H5Gclose(meta);
    const hsize_t chunk = 10000;  // limits.signal.chunksize default (parameters.cpp:615)
    hsize_t three = 3, two = 2, nf2[2] = {NF, 2};
    create_extendible(h, "qvectors", 2, &three, chunk);
    if (flags.fqt) {
        const hsize_t c2 = NF > 0 && NF < chunk ? NF : chunk;
        const hsize_t c1 = std::max<hsize_t>(1, chunk / c2);
        hsize_t dims[3] = {0, NF, 2}, mdims[3] = {H5S_UNLIMITED, NF, 2}, cdims[3] = {c1, c2, 2};
        const double fill = 0;
        hid_t dcpl = H5Pcreate(H5P_DATASET_CREATE);
        H5Pset_chunk(dcpl, 3, cdims);
        H5Pset_fill_value(dcpl, H5T_NATIVE_DOUBLE, &fill);
        hid_t space = H5Screate_simple(3, dims, mdims);
        hid_t ds = H5Dcreate2(h, "fqt", H5T_NATIVE_DOUBLE, space, H5P_DEFAULT, dcpl, H5P_DEFAULT);
        H5Sclose(space);
        H5Pclose(dcpl);
        if (ds < 0) throw terminate_request("cannot create dataset fqt");
        H5Dclose(ds);
    }
    (void)nf2;
    if (flags.fq0) create_extendible(h, "fq0", 2, &two, chunk);
    if (flags.fq) create_extendible(h, "fq", 2, &two, chunk);
    if (flags.fq2) create_extendible(h, "fq2", 2, &two, chunk);
    H5Fclose(h);
}

HDF5SignalWriter::~HDF5SignalWriter() {
    try {
        flush();
    } catch (...) {
    }
}

void HDF5SignalWriter::write(const CartesianCoor3D &q, const std::vector<std::complex<double>> &fqt,
                             std::complex<double> fq, std::complex<double> fq2) {
    q_.insert(q_.end(), {q.x, q.y, q.z});
    for (const auto &c : fqt) fqt_.insert(fqt_.end(), {c.real(), c.imag()});
    fq0_.insert(fq0_.end(), {fqt[0].real(), fqt[0].imag()});  // fq0 = fqt[0] (file_writer_service.cpp:514)
    fq_.insert(fq_.end(), {fq.real(), fq.imag()});
    fq2_.insert(fq2_.end(), {fq2.real(), fq2.imag()});
}

void HDF5SignalWriter::flush() {
    const size_t nq = q_.size() / 3;
    if (nq == 0) return;
    H5Quiet quiet;
    hid_t h = H5Fopen(file_.c_str(), H5F_ACC_RDWR, H5P_DEFAULT);
    if (h < 0) throw terminate_request("cannot open signal file: " + file_);
    append_rows(h, "qvectors", 2, q_, nq);
    append_rows(h, "fqt", 3, fqt_, nq);
    append_rows(h, "fq0", 2, fq0_, nq);
    append_rows(h, "fq", 2, fq_, nq);
    append_rows(h, "fq2", 2, fq2_, nq);
    H5Fclose(h);
    q_.clear();
    fqt_.clear();
    fq0_.clear();
    fq_.clear();
    fq2_.clear();
}

std::vector<CartesianCoor3D> HDF5SignalWriter::get_qvectors() const {
    H5Quiet quiet;
    std::vector<CartesianCoor3D> out;
    hid_t h = H5Fopen(file_.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT);
    if (h < 0) return out;
    hsize_t dims[2] = {0, 0};
    if (H5LTget_dataset_info(h, "qvectors", dims, nullptr, nullptr) >= 0 && dims[0] > 0) {
        std::vector<double> q(dims[0] * 3);
        H5LTread_dataset_double(h, "qvectors", q.data());
        for (size_t i = 0; i < dims[0]; ++i) out.emplace_back(q[3 * i], q[3 * i + 1], q[3 * i + 2]);
    }
    H5Fclose(h);
    return out;
}

std::string read_signal_meta(const std::string &h5, const std::string &name) {
    H5Quiet quiet;
    hid_t h = H5Fopen(h5.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT);
    if (h < 0) throw terminate_request("cannot open signal file: " + h5);
    const std::string path = "meta/" + name;
    std::string out;
    hid_t ds = H5Dopen2(h, path.c_str(), H5P_DEFAULT);
    if (ds < 0) {
        H5Fclose(h);
        throw terminate_request("signal file " + h5 + " has no " + path);
    }
    hid_t type = H5Dget_type(ds);
    if (H5Tget_class(type) == H5T_STRING) {
        const size_t n = H5Tget_size(type);
        std::vector<char> buf(n + 1, 0);
        hid_t mt = H5Tcopy(type);
        H5Dread(ds, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf.data());
        H5Tclose(mt);
        out.assign(buf.data());
    } else {
        hid_t space = H5Dget_space(ds);
        const hssize_t n = H5Sget_simple_extent_npoints(space);
        std::vector<char> buf((size_t)std::max<hssize_t>(n, 0));
        if (n > 0) H5Dread(ds, H5T_NATIVE_CHAR, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf.data());
        H5Sclose(space);
        out.assign(buf.begin(), buf.end());
    }
    H5Tclose(type);
    H5Dclose(ds);
    H5Fclose(h);
    return out;
}

void convert_signal_h5_to_raw(const std::string &h5, const std::string &raw) {
    H5Quiet quiet;
    hid_t h = H5Fopen(h5.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT);
    if (h < 0) throw terminate_request("cannot open " + h5);
    hsize_t qd[2] = {0, 0}, fd[3] = {0, 0, 0}, sd[2] = {0, 0};
    H5LTget_dataset_info(h, "qvectors", qd, nullptr, nullptr);
    H5LTget_dataset_info(h, "fqt", fd, nullptr, nullptr);
    const size_t nq = qd[0], NF = fd[1];
    std::vector<double> q(nq * 3), fqt(nq * NF * 2), fq(nq * 2), fq2(nq * 2);
    if (nq) {
        H5LTread_dataset_double(h, "qvectors", q.data());
        H5LTread_dataset_double(h, "fqt", fqt.data());
        if (H5LTget_dataset_info(h, "fq", sd, nullptr, nullptr) >= 0) H5LTread_dataset_double(h, "fq", fq.data());
        if (H5LTget_dataset_info(h, "fq2", sd, nullptr, nullptr) >= 0) H5LTread_dataset_double(h, "fq2", fq2.data());
    }
    H5Fclose(h);
    std::ofstream out(raw, std::ios::binary);
    uint64_t dims[2] = {nq, NF};
    out.write("SXSIG1", 6);
    out.write(reinterpret_cast<const char *>(dims), 16);
    for (size_t i = 0; i < nq; ++i) {
        out.write(reinterpret_cast<const char *>(&q[3 * i]), 24);
        out.write(reinterpret_cast<const char *>(&fqt[i * NF * 2]), sizeof(double) * 2 * NF);
        out.write(reinterpret_cast<const char *>(&fq[2 * i]), 16);
        out.write(reinterpret_cast<const char *>(&fq2[2 * i]), 16);
    }
}

// ------------------------------------------------------------ raw signal --

RawSignalWriter::RawSignalWriter(const std::string &file, size_t NF) : file_(file), NF_(NF) {
    std::ifstream in(file, std::ios::binary);
    if (in) {
        char magic[6];
        uint64_t dims[2];
        if (in.read(magic, 6) && !std::memcmp(magic, "SXSIG1", 6) && in.read(reinterpret_cast<char *>(dims), 16)) {
            if (dims[1] != NF)
                throw terminate_request("signal file " + file + " has " + std::to_string(dims[1]) + " frames");
            return;
        }
    }
    std::ofstream out(file, std::ios::binary | std::ios::trunc);
    uint64_t dims[2] = {0, NF};
    out.write("SXSIG1", 6);
    out.write(reinterpret_cast<const char *>(dims), 16);
}

void RawSignalWriter::write(const CartesianCoor3D &q, const std::vector<std::complex<double>> &fqt,
                            std::complex<double> fq, std::complex<double> fq2) {
    std::fstream io(file_, std::ios::binary | std::ios::in | std::ios::out);
    uint64_t dims[2];
    io.seekg(6);
    io.read(reinterpret_cast<char *>(dims), 16);
    io.seekp(0, std::ios::end);
    const double qq[3] = {q.x, q.y, q.z};
    io.write(reinterpret_cast<const char *>(qq), 24);
    io.write(reinterpret_cast<const char *>(fqt.data()), sizeof(double) * 2 * NF_);
    const double a[4] = {fq.real(), fq.imag(), fq2.real(), fq2.imag()};
    io.write(reinterpret_cast<const char *>(a), 32);
    dims[0] += 1;
    io.seekp(6);
    io.write(reinterpret_cast<const char *>(dims), 16);
}

std::vector<CartesianCoor3D> RawSignalWriter::get_qvectors() const {
    std::vector<CartesianCoor3D> out;
    std::ifstream in(file_, std::ios::binary);
    char magic[6];
    uint64_t dims[2];
    if (!in.read(magic, 6) || !in.read(reinterpret_cast<char *>(dims), 16)) return out;
    const long long rec = 24 + 16 * (long long)dims[1] + 32;
    for (uint64_t i = 0; i < dims[0]; ++i) {
        double q[3];
        in.seekg(22 + (long long)i * rec);
        in.read(reinterpret_cast<char *>(q), 24);
        out.emplace_back(q[0], q[1], q[2]);
    }
    return out;
}

std::unique_ptr<IResumableWriter> open_signal_writer(const RunFiles &files, size_t NF, const std::string &config,
                                                     const std::string &database_doc) {
    const std::string &f = files.signal_file;
    if (f.size() > 3 && f.compare(f.size() - 3, 3, ".h5") == 0) {
        auto slurp = [](const std::string &path) {
            std::ifstream in(path);
            std::stringstream ss;
            ss << in.rdbuf();
            return ss.str();
        };
        const std::string raw = config.empty() ? std::string() : slurp(config);
        const std::string db = files.database_file.empty() ? database_doc : dump_xml_text(slurp(files.database_file));
        return std::make_unique<HDF5SignalWriter>(f, NF, files, raw, dump_xml_text(raw), db);
    }
    return std::make_unique<RawSignalWriter>(f, NF);
}

std::vector<CartesianCoor3D> remaining_qvectors(const ISignalWriter &writer, const std::vector<CartesianCoor3D> &all) {
    // sassena.cpp:274-305: q vectors already present (exact match) are skipped
    const IResumableWriter *w = dynamic_cast<const IResumableWriter *>(&writer);
    std::set<std::tuple<double, double, double>> done;
    if (w)
        for (const auto &q : w->get_qvectors()) done.insert(std::make_tuple(q.x, q.y, q.z));
    std::vector<CartesianCoor3D> out;
    for (const auto &q : all)
        if (!done.count(std::make_tuple(q.x, q.y, q.z))) out.push_back(q);
    return out;
}

}  // namespace sassena_amd
