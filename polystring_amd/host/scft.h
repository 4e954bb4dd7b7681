// scft.h -- the SCFT driver (include/scft.h:56-109): scft(config, model) and
// run().  SINGLE cell runs relax() with the save_data stopping rules; the
// Brent / Manual cell scans are not part of this build.
#pragma once
#include <string>
#include <vector>

#include "Config.h"
#include "Model.h"

class scft {
public:
    scft(const std::string config_file, Model *pmodel);
    void run();
    int iterations() const { return _iter; }
    double last_H() const { return _H.empty() ? 0.0 : _H.back(); }
    double last_residual() const { return _residual_error.empty() ? 0.0 : _residual_error.back(); }

private:
    Config _cfg;
    Model *_pmodel;
    int _iter = 0;
    std::vector<double> _t, _time, _H, _residual_error, _incomp;
    double _curr_a = 0, _curr_b = 0, _curr_c = 0;

    double solve_scft(double a, double b, double c);
    void relax();
    bool save_data(const std::string s, const int i_save);
    void save_record(const std::string file);
    void display(const double t) const;
};
