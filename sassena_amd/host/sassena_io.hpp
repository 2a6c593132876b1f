// sassena_io.hpp -- input and output files around the hot path: the target's
// structure, trajectory readers (DCD, raw) and signal-file writers (the
// reference's HDF5 layout, and a raw layout), with resume support.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "sassena_config.hpp"
#include "sassena_host.hpp"

namespace sassena_amd {

// Element symbol of every target atom.  Formats: "elements" (one symbol per
// line) and "pdb" (ATOM/HETATM records; element columns 77-78, else the
// first letter of the atom name).
std::vector<std::string> read_structure(const std::string &file, const std::string &format);

// CHARMM/NAMD DCD trajectory (frames.cpp:272-436): Fortran records, optional
// unit-cell block before and fourth-dimension block after the x/y/z blocks.
class DCDTrajectory : public ICoordinateSource {
   public:
    explicit DCDTrajectory(const std::string &file);
    size_t atoms() const override { return NA_; }
    size_t frames() const override { return NF_; }
    void read_frame(size_t f, float *xyz) override;
    static bool detect(const std::string &file);

   private:
    std::string file_;
    size_t NA_ = 0, NF_ = 0;
    long long first_ = 0, block_ = 0, x_off_ = 0, y_off_ = 0, z_off_ = 0;
    std::vector<float> buf_;
};

// "SXRAW1" + uint64 NA, NF + float32 frames [NF][NA][3].
class RawTrajectory : public ICoordinateSource {
   public:
    explicit RawTrajectory(const std::string &file);
    size_t atoms() const override { return NA_; }
    size_t frames() const override { return NF_; }
    void read_frame(size_t f, float *xyz) override;
    static void write(const std::string &file, size_t NA, size_t NF, const float *frames);

   private:
    std::string file_;
    size_t NA_ = 0, NF_ = 0;
};

std::unique_ptr<ICoordinateSource> open_trajectory(const std::string &file, const std::string &format);

// DCD writer for tests and conversions (no unit cell).
void write_dcd(const std::string &file, size_t NA, size_t NF, const float *frames);

// A signal writer that also reports the q vectors already in its file.
class IResumableWriter : public ISignalWriter {
   public:
    virtual std::vector<CartesianCoor3D> get_qvectors() const = 0;
};

// The reference's signal file (file_writer_service.cpp:44-171, 314-484):
// qvectors [nq,3], fqt [nq,NF,2], fq0/fq/fq2 [nq,2] doubles, chunked and
// extendible, meta/{rawconfig,config,database}; appends to an existing file.
class HDF5SignalWriter : public IResumableWriter {
   public:
    // rawconfig: the configuration file's bytes (meta/rawconfig); config_doc
    // and database_doc: the parsed configuration and database documents
    // (meta/config, meta/database; file_writer_service.cpp:52-62)
    HDF5SignalWriter(const std::string &file, size_t NF, const RunFiles &flags, const std::string &rawconfig,
                     const std::string &config_doc, const std::string &database_doc);
    ~HDF5SignalWriter() override;
    void write(const CartesianCoor3D &q, const std::vector<std::complex<double>> &fqt, std::complex<double> fq,
               std::complex<double> fq2) override;
    void flush() override;
    std::vector<CartesianCoor3D> get_qvectors() const override;

   private:
    std::string file_;
    size_t NF_;
    RunFiles flags_;
    std::vector<double> q_, fqt_, fq0_, fq_, fq2_;
};

// "SXSIG1" + uint64 nq, NF + per shell q[3], fqt[NF][2], fq[2], fq2[2] doubles.
class RawSignalWriter : public IResumableWriter {
   public:
    RawSignalWriter(const std::string &file, size_t NF);
    void write(const CartesianCoor3D &q, const std::vector<std::complex<double>> &fqt, std::complex<double> fq,
               std::complex<double> fq2) override;
    std::vector<CartesianCoor3D> get_qvectors() const override;

   private:
    std::string file_;
    size_t NF_;
};

// config: path of the configuration file ("" for none); database_doc: the
// database document when no database file is named (files.database_file,
// whose document is stored otherwise).
std::unique_ptr<IResumableWriter> open_signal_writer(const RunFiles &files, size_t NF, const std::string &config,
                                                     const std::string &database_doc = std::string());
std::vector<CartesianCoor3D> remaining_qvectors(const ISignalWriter &writer, const std::vector<CartesianCoor3D> &all);

// HDF5 signal file -> raw signal layout (for tools without an HDF5 reader).
void convert_signal_h5_to_raw(const std::string &h5, const std::string &raw);
// The meta/<name> dataset of a signal file (string or character array).
std::string read_signal_meta(const std::string &h5, const std::string &name);

}  // namespace sassena_amd
