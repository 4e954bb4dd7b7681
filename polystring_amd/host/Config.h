// Config.h -- the param.ini reader of the AB / periodic hot path.
//
// Same class name, constructor and accessor names as the reference's Config
// (include/Config.h:51-305) for the keys Model_AB and scft read on this path;
// the INI syntax is SimpleIni's: [Section] headers, "key = value" lines, '#'
// or ';' comments, Python-style lists "[a, b, c]", case-insensitive enum
// words (src/Config.cc:1099-1110, 1219-1233, 1495).  Extension keys read by
// this build only: [GPU] device (default 0) and [GPU] precision
// (FP64 | FP32, default FP64).
#pragma once
#include <map>
#include <string>
#include <vector>

#include "common.h"

class Config {
public:
    Config() {}
    explicit Config(const char *config_file);
    explicit Config(const std::string &config_file) : Config(config_file.c_str()) {}

    bool check() const;

    // generic access
    std::string get_string(const std::string &section, const std::string &key) const;
    double get_double(const std::string &section, const std::string &key, double dflt = 0.0) const;
    int get_int(const std::string &section, const std::string &key, int dflt = 0) const;
    bool get_bool(const std::string &section, const std::string &key, bool dflt = false) const;
    std::vector<double> get_list(const std::string &section, const std::string &key) const;
    void set(const std::string &section, const std::string &key, const std::string &value);

    // [IO]
    std::string base_dir() const;
    std::string data_file() const;
    std::string param_file() const;
    bool is_display() const;
    bool is_save_data() const;
    bool is_save_q() const;
    unsigned display_interval() const;
    unsigned record_interval() const;
    unsigned save_interval() const;
    // [Model]
    ModelType model() const;
    std::string get_model_type_string() const;
    std::vector<double> f() const;
    std::vector<double> chiN() const;
    std::vector<double> segment_length() const;
    bool is_compressible() const;
    // [UnitCell]
    CrystalSystemType get_crystal_system_type() const;
    std::string get_crystal_system_type_string() const;
    double a() const;
    double b() const;
    double c() const;
    double a(double v);
    double b(double v);
    double c(double v);
    double alpha() const;
    double beta() const;
    double gamma() const;
    // [Grid]
    unsigned dim() const;
    unsigned Lx() const;
    unsigned Ly() const;
    unsigned Lz() const;
    GridInitType get_grid_init_type() const;
    std::string field_data_file() const;
    unsigned seed() const;
    // [Algorithm_MDE]
    AlgorithmMDEType algo_mde_type() const;
    std::string get_algo_mde_type_string() const;
    std::vector<double> ds() const;
    std::vector<unsigned> Ms() const;
    int etdrk4_M() const;  // src/Config.cc:1203-1205 (0 when absent)
    // [Algorithm_SCFT]
    AlgorithmSCFTType algo_scft_type() const;
    std::vector<double> lam() const;
    unsigned min_iter() const;
    unsigned max_iter() const;
    double thresh_H() const;
    double thresh_residual() const;
    unsigned n_Anderson_mixing() const;
    // [Algorithm_Cell_Optimization]
    AlgorithmCellType algo_cell_optimization_type() const;
    // [Algorithm_Contour_Integration]
    AlgorithmContourType algo_contour_integration_type() const;
    // [GPU] (extension)
    int device() const;
    int precision() const;  // 0 FP64, 1 FP32 bulk

private:
    std::map<std::string, std::map<std::string, std::string>> _kv;  // lower-case section -> key -> value
    std::string _file;
};
