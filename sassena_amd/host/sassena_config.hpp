// sassena_config.hpp -- the scatter.xml subset of the hot path.
//
// A small XML element tree (elements, text and attributes; comments,
// processing instructions and CDATA are skipped) and the reader of the keys
// Params::read_xml takes for scattering, stager and sample input
// (parameters.cpp:345-606).  Relative file names resolve against the
// directory of the configuration file (parameters.cpp:903-908).
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "sassena_host.hpp"

namespace sassena_amd {

struct XMLElement {
    std::string name;
    std::string text;
    std::vector<std::pair<std::string, std::string>> attributes;
    std::vector<std::unique_ptr<XMLElement>> children;
    // first descendant on a '/'-separated relative path, or nullptr
    const XMLElement *find(const std::string &path) const;
    std::vector<const XMLElement *> find_all(const std::string &path) const;
    bool exists(const std::string &path) const { return find(path) != nullptr; }
    std::string get(const std::string &path, const std::string &def) const;
    double get(const std::string &path, double def) const;
    long get(const std::string &path, long def) const;
    bool get(const std::string &path, bool def) const;
};

std::unique_ptr<XMLElement> parse_xml(const std::string &text);
std::unique_ptr<XMLElement> parse_xml_file(const std::string &path);
// The parsed document serialised again (declaration, elements, attributes
// and text; comments and processing instructions dropped, two-space
// indentation): what the reference stores as meta/config and meta/database
// through libxml2's document dump (xml_interface.cpp:121-128,
// file_writer_service.cpp:58-62).  dump_xml_text: "" when the text does not parse.
std::string dump_xml(const XMLElement &root);
std::string dump_xml_text(const std::string &text);

// Input and output files named in the configuration.
struct RunFiles {
    std::string structure_file, structure_format = "elements";  // sample.structure
    std::string frames_file, frames_format = "raw";             // sample.framesets.frameset
    std::string signal_file = "signal.h5";                      // scattering.signal.file
    std::string database_file;  // database.file when given; else db.xml beside the configuration if present
    bool fqt = true, fq0 = true, fq = true, fq2 = true;
};

// Reads the configuration (root element <root>); fills params (generating the
// q vectors and orientation vectors as the reference does) and files.
void read_config(const std::string &path, Params &params, RunFiles &files);
void read_config_text(const std::string &text, const std::string &basedir, Params &params, RunFiles &files);

// Database::read_xml (database.cpp:31-145): elements of //scatterfactors
// (required), //sizes and //exclusionfactors, each <element> with <name>,
// <type> and a list of <param>.  Masses and PDB name patterns are not used
// on this path and are skipped.
Database read_database_xml(const std::string &path);
Database read_database_xml_text(const std::string &text);

}  // namespace sassena_amd
