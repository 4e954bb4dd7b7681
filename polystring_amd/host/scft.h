// scft.h -- the SCFT driver (include/scft.h:56-109): scft(config, model) and
// run().  SINGLE runs relax() with the save_data stopping rules; MANUAL /
// BATCH scans the batch_cell_min/step/max sizes as the reference does; BRENT
// runs the stress-driven cell relaxation that replaces Brent's search.
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

    std::vector<double> _as, _bs, _cs, _Fs;  // cell record (save_cell, src/scft.cc:589-609)

    double solve_scft(double a, double b, double c);
    void reset_model(double a, double b, double c);
    void save_cell(const std::string file) const;
    void relax();
    bool save_data(const std::string s, const int i_save);
    void save_record(const std::string file);
    void display(const double t) const;
};
