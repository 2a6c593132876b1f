// sassena_config.cpp -- XML reader and the scatter.xml hot-path subset.
#include "sassena_config.hpp"

#include <cctype>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace sassena_amd {

// ------------------------------------------------------------ XML ---------

namespace {

struct Parser {
    const std::string &s;
    size_t i = 0;
    explicit Parser(const std::string &t) : s(t) {}

    [[noreturn]] void error(const std::string &msg) const {
        throw terminate_request("XML parse error at offset " + std::to_string(i) + ": " + msg);
    }
    bool starts(const char *p) const { return s.compare(i, std::char_traits<char>::length(p), p) == 0; }
    void skip_ws() {
        while (i < s.size() && std::isspace((unsigned char)s[i])) ++i;
    }
    void skip_until(const char *end) {
        const size_t k = s.find(end, i);
        if (k == std::string::npos) error(std::string("unterminated construct, expected ") + end);
        i = k + std::char_traits<char>::length(end);
    }
    // skips comments, processing instructions and doctype declarations
    void skip_misc() {
        for (;;) {
            skip_ws();
            if (starts("<!--")) skip_until("-->");
            else if (starts("<?")) skip_until("?>");
            else if (starts("<!DOCTYPE")) skip_until(">");
            else return;
        }
    }
    std::string name() {
        const size_t b = i;
        while (i < s.size() && (std::isalnum((unsigned char)s[i]) || s[i] == '_' || s[i] == '-' || s[i] == '.' ||
                                s[i] == ':'))
            ++i;
        if (i == b) error("expected a name");
        return s.substr(b, i - b);
    }
    static std::string decode(const std::string &t) {
        std::string out;
        for (size_t k = 0; k < t.size(); ++k) {
            if (t[k] == '&') {
                const size_t e = t.find(';', k);
                if (e != std::string::npos) {
                    const std::string ent = t.substr(k + 1, e - k - 1);
                    if (ent == "lt") out += '<';
                    else if (ent == "gt") out += '>';
                    else if (ent == "amp") out += '&';
                    else if (ent == "quot") out += '"';
                    else if (ent == "apos") out += '\'';
                    else out += t.substr(k, e - k + 1);
                    k = e;
                    continue;
                }
            }
            out += t[k];
        }
        return out;
    }
    std::unique_ptr<XMLElement> element() {
        if (i >= s.size() || s[i] != '<') error("expected '<'");
        ++i;
        auto el = std::make_unique<XMLElement>();
        el->name = name();
        for (;;) {
            skip_ws();
            if (starts("/>")) {
                i += 2;
                return el;
            }
            if (i < s.size() && s[i] == '>') {
                ++i;
                break;
            }
            const std::string an = name();
            skip_ws();
            if (i >= s.size() || s[i] != '=') error("expected '=' after attribute");
            ++i;
            skip_ws();
            if (i >= s.size() || (s[i] != '"' && s[i] != '\'')) error("expected quoted attribute value");
            const char q = s[i++];
            const size_t e = s.find(q, i);
            if (e == std::string::npos) error("unterminated attribute value");
            el->attributes.emplace_back(an, decode(s.substr(i, e - i)));
            i = e + 1;
        }
        std::string text;
        for (;;) {
            if (i >= s.size()) error("unterminated element <" + el->name + ">");
            if (starts("<!--")) {
                skip_until("-->");
            } else if (starts("<![CDATA[")) {
                i += 9;
                const size_t e = s.find("]]>", i);
                if (e == std::string::npos) error("unterminated CDATA");
                text += s.substr(i, e - i);
                i = e + 3;
            } else if (starts("<?")) {
                skip_until("?>");
            } else if (starts("</")) {
                i += 2;
                const std::string n = name();
                if (n != el->name) error("mismatched </" + n + "> for <" + el->name + ">");
                skip_ws();
                if (i >= s.size() || s[i] != '>') error("expected '>'");
                ++i;
                break;
            } else if (s[i] == '<') {
                el->children.push_back(element());
            } else {
                const size_t e = s.find('<', i);
                text += s.substr(i, (e == std::string::npos ? s.size() : e) - i);
                i = e == std::string::npos ? s.size() : e;
            }
        }
        // trimmed text content
        size_t b = 0, e = text.size();
        while (b < e && std::isspace((unsigned char)text[b])) ++b;
        while (e > b && std::isspace((unsigned char)text[e - 1])) --e;
        el->text = decode(text.substr(b, e - b));
        return el;
    }
};

void split_path(const std::string &path, std::vector<std::string> &parts) {
    std::stringstream ss(path);
    std::string p;
    while (std::getline(ss, p, '/'))
        if (!p.empty() && p != ".") parts.push_back(p);
}

void collect(const XMLElement *e, const std::vector<std::string> &parts, size_t k,
             std::vector<const XMLElement *> &out) {
    if (k == parts.size()) {
        out.push_back(e);
        return;
    }
    for (const auto &c : e->children)
        if (c->name == parts[k]) collect(c.get(), parts, k + 1, out);
}

std::string encode(const std::string &t) {
    std::string out;
    for (char ch : t) {
        switch (ch) {
            case '<': out += "&lt;"; break;
            case '>': out += "&gt;"; break;
            case '&': out += "&amp;"; break;
            case '"': out += "&quot;"; break;
            default: out += ch;
        }
    }
    return out;
}

void dump_element(const XMLElement &e, int depth, std::string &out) {
    out.append(2 * depth, ' ');
    out += "<" + e.name;
    for (const auto &a : e.attributes) out += " " + a.first + "=\"" + encode(a.second) + "\"";
    if (e.children.empty() && e.text.empty()) {
        out += "/>\n";
        return;
    }
    out += ">";
    if (e.children.empty()) {
        out += encode(e.text) + "</" + e.name + ">\n";
        return;
    }
    out += "\n";
    if (!e.text.empty()) {
        out.append(2 * depth + 2, ' ');
        out += encode(e.text) + "\n";
    }
    for (const auto &c : e.children) dump_element(*c, depth + 1, out);
    out.append(2 * depth, ' ');
    out += "</" + e.name + ">\n";
}

}  // namespace

std::string dump_xml(const XMLElement &root) {
    std::string out = "<?xml version=\"1.0\"?>\n";
    dump_element(root, 0, out);
    return out;
}

std::string dump_xml_text(const std::string &text) {
    try {
        return dump_xml(*parse_xml(text));
    } catch (const terminate_request &) {
        return std::string();
    }
}

std::vector<const XMLElement *> XMLElement::find_all(const std::string &path) const {
    std::vector<std::string> parts;
    split_path(path, parts);
    std::vector<const XMLElement *> out;
    collect(this, parts, 0, out);
    return out;
}

const XMLElement *XMLElement::find(const std::string &path) const {
    std::vector<const XMLElement *> all = find_all(path);
    return all.empty() ? nullptr : all[0];
}

std::string XMLElement::get(const std::string &path, const std::string &def) const {
    const XMLElement *e = find(path);
    return e ? e->text : def;
}

double XMLElement::get(const std::string &path, double def) const {
    const XMLElement *e = find(path);
    if (!e) return def;
    char *end = nullptr;
    const double v = std::strtod(e->text.c_str(), &end);
    if (end == e->text.c_str()) throw terminate_request("not a number at " + path + ": '" + e->text + "'");
    return v;
}

long XMLElement::get(const std::string &path, long def) const {
    const XMLElement *e = find(path);
    if (!e) return def;
    char *end = nullptr;
    const long v = std::strtol(e->text.c_str(), &end, 10);
    if (end == e->text.c_str()) throw terminate_request("not an integer at " + path + ": '" + e->text + "'");
    return v;
}

bool XMLElement::get(const std::string &path, bool def) const {
    const XMLElement *e = find(path);
    if (!e) return def;
    const std::string &t = e->text;
    if (t == "true" || t == "1" || t == "yes") return true;
    if (t == "false" || t == "0" || t == "no") return false;
    throw terminate_request("not a boolean at " + path + ": '" + t + "'");
}

std::unique_ptr<XMLElement> parse_xml(const std::string &text) {
    Parser p(text);
    p.skip_misc();
    std::unique_ptr<XMLElement> root = p.element();
    p.skip_misc();
    if (p.i != text.size()) p.error("trailing content after the root element");
    return root;
}

std::unique_ptr<XMLElement> parse_xml_file(const std::string &path) {
    std::ifstream in(path);
    if (!in) throw terminate_request("cannot open configuration file: " + path);
    std::stringstream ss;
    ss << in.rdbuf();
    return parse_xml(ss.str());
}

// ------------------------------------------------------------ scatter.xml -

static std::string resolve(const std::string &basedir, const std::string &file) {
    if (file.empty() || file[0] == '/' || basedir.empty()) return file;
    return basedir + "/" + file;
}

static CartesianCoor3D read_vec(const XMLElement *e, CartesianCoor3D def) {
    if (!e) return def;
    return {e->get("x", def.x), e->get("y", def.y), e->get("z", def.z)};
}

Database read_database_xml_text(const std::string &text) {
    std::unique_ptr<XMLElement> root = parse_xml(text);
    struct Entry {
        bool has_sf = false;
        ElementEntry e;
    };
    std::vector<std::string> names;
    std::vector<Entry> entries;
    auto slot = [&](const std::string &name) -> Entry & {
        for (size_t i = 0; i < names.size(); ++i)
            if (names[i] == name) return entries[i];
        names.push_back(name);
        entries.emplace_back();
        return entries.back();
    };
    auto read_function = [](const XMLElement *el, DatabaseFunction &f) {
        f.function_type = (int)el->get("type", 0L);
        f.constants.clear();
        for (const XMLElement *p : el->find_all("param")) {
            char *end = nullptr;
            const double v = std::strtod(p->text.c_str(), &end);
            if (end == p->text.c_str()) throw terminate_request("database: bad <param> value '" + p->text + "'");
            f.constants.push_back(v);
        }
    };
    const XMLElement *sf = root->name == "scatterfactors" ? root.get() : root->find("scatterfactors");
    if (!sf) throw terminate_request("Need scattering factors.");
    for (const XMLElement *el : sf->find_all("element")) {
        Entry &e = slot(el->get("name", std::string()));
        read_function(el, e.e.sfactor);
        e.has_sf = true;
    }
    if (const XMLElement *sz = root->find("sizes"))
        for (const XMLElement *el : sz->find_all("element")) read_function(el, slot(el->get("name", std::string())).e.volume);
    if (const XMLElement *ex = root->find("exclusionfactors"))
        for (const XMLElement *el : ex->find_all("element"))
            read_function(el, slot(el->get("name", std::string())).e.exclusion);
    Database db;
    for (size_t i = 0; i < names.size(); ++i) {
        if (names[i].empty()) throw terminate_request("database: <element> without <name>");
        if (!entries[i].has_sf) continue;  // only elements with a scattering factor can be looked up
        db.add(names[i], entries[i].e);
    }
    return db;
}

Database read_database_xml(const std::string &path) {
    std::ifstream in(path);
    if (!in) throw terminate_request("cannot open database file: " + path);
    std::stringstream ss;
    ss << in.rdbuf();
    return read_database_xml_text(ss.str());
}

void read_config_text(const std::string &text, const std::string &basedir, Params &params, RunFiles &files) {
    std::unique_ptr<XMLElement> root = parse_xml(text);
    const XMLElement &x = *root;
    ScatteringParameters &sc = params.scattering;

    // database (parameters.cpp:775-800): type file, format xml
    if (const XMLElement *d = x.find("database")) {
        const std::string type = d->get("type", std::string("file"));
        const std::string format = d->get("format", std::string("xml"));
        if (type != "file") throw terminate_request("database.type not understood: " + type);
        if (format != "xml") throw terminate_request("database.format not understood: " + format);
        files.database_file = resolve(basedir, d->get("file", std::string("db.xml")));
    }

    // sample (structure and trajectory of the target)
    files.structure_file = resolve(basedir, x.get("sample/structure/file", std::string()));
    files.structure_format = x.get("sample/structure/format", files.structure_format);
    files.frames_file = resolve(basedir, x.get("sample/framesets/frameset/file", std::string()));
    files.frames_format = x.get("sample/framesets/frameset/format", files.frames_format);
    params.stager_target = x.get("stager/target", params.stager_target);
    params.device = (int)x.get("limits/computation/device", (long)params.device);
    params.batch_shells = (size_t)x.get("limits/computation/shells", (long)params.batch_shells);
    params.limits.stage_data = (size_t)x.get("limits/stage/memory/data", (long)params.limits.stage_data);
    params.limits.result_buffer =
        (size_t)x.get("limits/computation/memory/result_buffer", (long)params.limits.result_buffer);
    params.limits.scale = (size_t)x.get("limits/computation/memory/scale", (long)params.limits.scale);

    // scattering (parameters.cpp:370-606)
    if (const XMLElement *s = x.find("scattering")) {
        if (s->exists("target")) throw terminate_request("scattering.target is obsolete. Use stager.target instead.");
        sc.type = s->get("type", sc.type);
        sc.background_factor = s->get("background/factor", sc.background_factor);
        sc.dsp_type = s->get("dsp/type", sc.dsp_type);
        sc.dsp_method = s->get("dsp/method", sc.dsp_method);
        // q vectors: single | scans | file
        if (s->exists("vectors")) {
            const std::string vt = s->get("vectors/type", std::string("single"));
            if (vt == "single") {
                sc.qvectors.push_back(read_vec(s->find("vectors/single"), {1.0, 0.0, 0.0}));
            } else if (vt == "scans") {
                for (const XMLElement *e : s->find_all("vectors/scans/scan")) {
                    ScanParameters scan;
                    scan.from = 0;
                    scan.to = 1;
                    scan.basevector = read_vec(e->find("base"), {1.0, 0.0, 0.0});
                    scan.from = e->get("from", scan.from);
                    scan.to = e->get("to", scan.to);
                    scan.points = (size_t)e->get("points", (long)scan.points);
                    scan.exponent = e->get("exponent", scan.exponent);
                    sc.scans.push_back(scan);
                }
                sc.create_from_scans();
            } else if (vt == "file") {
                const std::string f = resolve(basedir, s->get("vectors/file", std::string("qvectors.txt")));
                std::ifstream in(f);
                if (!in) throw terminate_request("cannot open q vector file: " + f);
                double a, b, c;
                while (in >> a >> b >> c) sc.qvectors.emplace_back(a, b, c);
            } else {
                throw terminate_request("scattering.vectors.type not understood: " + vt);
            }
            if (sc.qvectors.empty())
                throw terminate_request("No q vectors generated. Check the scattering.vectors section for errors.");
        }
        // orientational averaging
        if (const XMLElement *o = s->find("average/orientation")) {
            sc.axis = read_vec(o->find("axis"), sc.axis);
            sc.orientation_type = o->get("type", sc.orientation_type);
            if (sc.orientation_type == "vectors") {
                OrientationVectors &v = sc.vectors;
                v.type = o->get("vectors/type", v.type);
                v.algorithm = o->get("vectors/algorithm", v.algorithm);
                v.resolution = o->get("vectors/resolution", v.resolution);
                v.seed = (unsigned long)o->get("vectors/seed", (long)v.seed);
                v.filepath = resolve(basedir, o->get("vectors/file", v.filepath));
                v.create();
            } else if (sc.orientation_type == "multipole") {
                sc.multipole_type = o->get("multipole/type", sc.multipole_type);
                MultipoleMoments &mm = sc.moments;
                mm.type = o->get("multipole/moments/type", mm.type);
                mm.resolution = o->get("multipole/moments/resolution", mm.resolution);
                mm.filepath = resolve(basedir, o->get("multipole/moments/file", mm.filepath));
                mm.create(sc.multipole_type);
            } else if (sc.orientation_type != "none") {
                throw terminate_request("Orientation averaging type not understood: " + sc.orientation_type);
            }
        }
        files.signal_file = resolve(basedir, s->get("signal/file", files.signal_file));
        files.fqt = s->get("signal/fqt", files.fqt);
        files.fq0 = s->get("signal/fq0", files.fq0);
        files.fq = s->get("signal/fq", files.fq);
        files.fq2 = s->get("signal/fq2", files.fq2);
    }
}

void read_config(const std::string &path, Params &params, RunFiles &files) {
    std::ifstream in(path);
    if (!in) throw terminate_request("cannot open configuration file: " + path);
    std::stringstream ss;
    ss << in.rdbuf();
    const size_t slash = path.find_last_of('/');
    const std::string basedir = slash == std::string::npos ? std::string() : path.substr(0, slash);
    read_config_text(ss.str(), basedir, params, files);
}

}  // namespace sassena_amd
