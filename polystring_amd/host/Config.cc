// Config.cc -- INI parsing and the hot-path accessors (see Config.h).
#include "Config.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <sstream>

namespace {

std::string trim(const std::string &s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) ++a;
    while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
    return s.substr(a, b - a);
}

std::string lower(std::string s) {
    std::transform(s.begin(), s.end(), s.begin(), [](unsigned char c) { return (char)std::tolower(c); });
    return s;
}

std::string upper(std::string s) {
    std::transform(s.begin(), s.end(), s.begin(), [](unsigned char c) { return (char)std::toupper(c); });
    return s;
}

}  // namespace

Config::Config(const char *config_file) : _file(config_file) {
    std::ifstream in(config_file);
    if (!in) {
        std::cerr << "Config: cannot open " << config_file << std::endl;
        std::exit(1);
    }
    std::string line, section;
    while (std::getline(in, line)) {
        line = trim(line);
        if (line.empty() || line[0] == '#' || line[0] == ';') continue;
        if (line[0] == '[' && line.back() == ']') {
            section = lower(trim(line.substr(1, line.size() - 2)));
            continue;
        }
        const size_t eq = line.find('=');
        if (eq == std::string::npos) continue;
        _kv[section][lower(trim(line.substr(0, eq)))] = trim(line.substr(eq + 1));
    }
    if (!check()) std::exit(1);
}

std::string Config::get_string(const std::string &section, const std::string &key) const {
    auto s = _kv.find(lower(section));
    if (s == _kv.end()) return "";
    auto k = s->second.find(lower(key));
    return k == s->second.end() ? "" : k->second;
}

void Config::set(const std::string &section, const std::string &key, const std::string &value) {
    _kv[lower(section)][lower(key)] = value;
}

double Config::get_double(const std::string &section, const std::string &key, double dflt) const {
    const std::string v = get_string(section, key);
    return v.empty() ? dflt : std::strtod(v.c_str(), nullptr);
}

int Config::get_int(const std::string &section, const std::string &key, int dflt) const {
    const std::string v = get_string(section, key);
    return v.empty() ? dflt : (int)std::strtol(v.c_str(), nullptr, 10);
}

bool Config::get_bool(const std::string &section, const std::string &key, bool dflt) const {
    const std::string v = lower(get_string(section, key));
    if (v.empty()) return dflt;
    return v == "true" || v == "1" || v == "yes" || v == "on";
}

// Python-style list "[1.0, 2, 3e-1]" (src/Config.cc:1549-1624)
std::vector<double> Config::get_list(const std::string &section, const std::string &key) const {
    std::string v = get_string(section, key);
    std::vector<double> out;
    for (char &ch : v)
        if (ch == '[' || ch == ']' || ch == ',') ch = ' ';
    std::istringstream is(v);
    double d;
    while (is >> d) out.push_back(d);
    return out;
}

bool Config::check() const {
    // Anderson mixing requires File initialisation (src/Config.cc:1304-1320)
    if (algo_scft_type() == AlgorithmSCFTType::ANDERSON && get_grid_init_type() != GridInitType::FILE_INIT) {
        std::cerr << "Anderson mixing requires gridInitType = File." << std::endl;
        return false;
    }
    if (dim() < 1 || dim() > 3) {
        std::cerr << "Grid dimension must be 1, 2 or 3." << std::endl;
        return false;
    }
    if (Ms().size() < 2) {
        std::cerr << "[Algorithm_MDE] Ms must list two slice counts." << std::endl;
        return false;
    }
    return true;
}

std::string Config::base_dir() const {
    std::string d = get_string("IO", "base_dir");
    return d.empty() ? "./" : d;
}
std::string Config::data_file() const { return get_string("IO", "data_file"); }
std::string Config::param_file() const { return get_string("IO", "param_file"); }
bool Config::is_display() const { return get_bool("IO", "is_display"); }
bool Config::is_save_data() const { return get_bool("IO", "is_save_data"); }
bool Config::is_save_q() const { return get_bool("IO", "is_save_q"); }
unsigned Config::display_interval() const { return (unsigned)std::max(1, get_int("IO", "display_interval", 1)); }
unsigned Config::record_interval() const { return (unsigned)std::max(1, get_int("IO", "record_interval", 1)); }
unsigned Config::save_interval() const { return (unsigned)std::max(1, get_int("IO", "save_interval", 1000000)); }

ModelType Config::model() const { return upper(get_string("Model", "model")) == "AB" ? ModelType::AB : ModelType::UNKNOWN; }
std::string Config::get_model_type_string() const { return get_string("Model", "model"); }
std::vector<double> Config::f() const { return get_list("Model", "f"); }
std::vector<double> Config::chiN() const { return get_list("Model", "chiN"); }
std::vector<double> Config::segment_length() const { return get_list("Model", "a"); }
bool Config::is_compressible() const { return get_bool("Model", "is_compressible"); }

CrystalSystemType Config::get_crystal_system_type() const {
    const std::string s = upper(get_string("UnitCell", "CrystalSystemType"));
    if (s == "LAM" || s == "LAMELLAR" || s == "LAMELLA" || s == "LAMELLAE") return CrystalSystemType::LAMELLAR;
    if (s == "SQUARE" || s == "SQ") return CrystalSystemType::SQUARE;
    if (s == "RECT" || s == "RECTANGLE" || s == "RECTANGULAR") return CrystalSystemType::RECTANGULAR;
    if (s == "HEX" || s == "HEXAGONAL" || s == "HEXAGON") return CrystalSystemType::HEXAGONAL;
    if (s == "OB" || s == "OBLIQUE") return CrystalSystemType::OBLIQUE;
    if (s == "CUBE" || s == "CUBIC") return CrystalSystemType::CUBIC;
    if (s == "TETRAGONAL" || s == "TETRAGON") return CrystalSystemType::TETRAGONAL;
    if (s == "ORTHORHOMBIC") return CrystalSystemType::ORTHORHOMBIC;
    if (s == "TRIGONAL") return CrystalSystemType::TRIGONAL;
    if (s == "MONOCLINIC") return CrystalSystemType::MONOCLINIC;
    if (s == "TRICLINIC") return CrystalSystemType::TRICLINIC;
    return CrystalSystemType::UNKNOWN;
}
std::string Config::get_crystal_system_type_string() const { return get_string("UnitCell", "CrystalSystemType"); }
double Config::a() const { return get_double("UnitCell", "a"); }
double Config::b() const { return get_double("UnitCell", "b"); }
double Config::c() const { return get_double("UnitCell", "c"); }
double Config::a(double v) { std::ostringstream s; s.precision(17); s << v; set("UnitCell", "a", s.str()); return v; }
double Config::b(double v) { std::ostringstream s; s.precision(17); s << v; set("UnitCell", "b", s.str()); return v; }
double Config::c(double v) { std::ostringstream s; s.precision(17); s << v; set("UnitCell", "c", s.str()); return v; }
// angles are given in degrees in param.ini and returned in radians
double Config::alpha() const { return (PI / 180.0) * get_double("UnitCell", "alpha"); }
double Config::beta() const { return (PI / 180.0) * get_double("UnitCell", "beta"); }
double Config::gamma() const { return (PI / 180.0) * get_double("UnitCell", "gamma"); }

unsigned Config::dim() const { return (unsigned)get_int("Grid", "dimension", 3); }
unsigned Config::Lx() const { return (unsigned)get_int("Grid", "Lx", 1); }
unsigned Config::Ly() const { return (unsigned)get_int("Grid", "Ly", 1); }
unsigned Config::Lz() const { return (unsigned)get_int("Grid", "Lz", 1); }
GridInitType Config::get_grid_init_type() const {
    const std::string s = upper(get_string("Grid", "gridInitType"));
    if (s == "FILE") return GridInitType::FILE_INIT;
    if (s == "RAN" || s == "RANDOM") return GridInitType::RANDOM_INIT;
    if (s == "CONST" || s == "CONSTANT") return GridInitType::CONSTANT_INIT;
    if (s == "DATA") return GridInitType::DATA_INIT;
    if (s == "PATTERN") return GridInitType::PATTERN_INIT;
    return GridInitType::UNKNOWN;
}
std::string Config::field_data_file() const { return get_string("Grid", "field_data"); }
unsigned Config::seed() const { return (unsigned)get_int("Grid", "random_seed", 0); }

AlgorithmMDEType Config::algo_mde_type() const {
    const std::string s = upper(get_string("Algorithm_MDE", "algorithm"));
    if (s == "PSEUDOSPECTRAL" || s == "OS2" || s == "OS") return AlgorithmMDEType::OS2;
    if (s == "RQM4" || s == "OS4") return AlgorithmMDEType::RQM4;
    return AlgorithmMDEType::ETDRK4;
}
std::string Config::get_algo_mde_type_string() const { return get_string("Algorithm_MDE", "algorithm"); }
std::vector<double> Config::ds() const { return get_list("Algorithm_MDE", "ds"); }
int Config::etdrk4_M() const { return get_int("Algorithm_MDE", "etdrk4_M", 0); }

std::vector<unsigned> Config::Ms() const {
    std::vector<unsigned> out;
    for (double d : get_list("Algorithm_MDE", "Ms")) out.push_back((unsigned)std::lround(d));
    return out;
}

AlgorithmSCFTType Config::algo_scft_type() const {
    const std::string s = upper(get_string("Algorithm_SCFT", "algorithm"));
    return s.rfind("ANDERSON", 0) == 0 ? AlgorithmSCFTType::ANDERSON : AlgorithmSCFTType::EM;
}
std::vector<double> Config::lam() const { return get_list("Algorithm_SCFT", "lam"); }
unsigned Config::min_iter() const { return (unsigned)get_int("Algorithm_SCFT", "min_iter", 10); }
unsigned Config::max_iter() const { return (unsigned)get_int("Algorithm_SCFT", "max_iter", 1000); }
double Config::thresh_H() const { return get_double("Algorithm_SCFT", "thresh_H", 1e-6); }
double Config::thresh_residual() const { return get_double("Algorithm_SCFT", "thresh_residual", 1e-8); }
unsigned Config::n_Anderson_mixing() const { return (unsigned)get_int("Algorithm_SCFT", "n_Anderson_mixing", 20); }

AlgorithmCellType Config::algo_cell_optimization_type() const {
    const std::string s = upper(get_string("Algorithm_Cell_Optimization", "algorithm"));
    // src/Config.cc:1334-1345
    if (s == "BRENT" || s == "BRENT METHOD" || s == "NEWTON") return AlgorithmCellType::BRENT;
    if (s == "MANUAL" || s == "BATCH") return AlgorithmCellType::MANUAL;
    return AlgorithmCellType::SINGLE;
}

AlgorithmContourType Config::algo_contour_integration_type() const {
    const std::string s = upper(get_string("Algorithm_Contour_Integration", "algorithm"));
    return s == "TRAPEZOIDAL" ? AlgorithmContourType::TRAPEZOIDAL : AlgorithmContourType::SIMPSON;
}

int Config::device() const { return get_int("GPU", "device", 0); }
int Config::precision() const { return upper(get_string("GPU", "precision")) == "FP32" ? 1 : 0; }
