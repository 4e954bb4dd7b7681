// scft.cc -- relax() and save_data() of src/scft.cc:382-422,532-587.
#include "scft.h"

#include <chrono>
#include <cstdlib>
#include <iostream>
#include <sstream>

#include "UnitCell.h"
#include "mat5.h"

scft::scft(const std::string config_file, Model *pmodel) : _cfg(config_file.c_str()), _pmodel(pmodel) {
    UnitCell uc(_cfg);
    _curr_a = uc.lx();
    _curr_b = uc.ly();
    _curr_c = uc.lz();
}

void scft::run() {
    if (_cfg.algo_cell_optimization_type() != AlgorithmCellType::SINGLE) {
        std::cerr << "Cell optimization algorithm other than Single is not available in this build." << std::endl;
        std::exit(1);
    }
    solve_scft(_cfg.a(), _cfg.b(), _cfg.c());
}

double scft::solve_scft(double a, double b, double c) {
    (void)a; (void)b; (void)c;
    relax();
    return _H.empty() ? 0.0 : _H.back();
}

void scft::display(const double t) const {
    std::cout << _iter << std::endl;
    std::cout << "t = " << t << std::endl;
    _pmodel->display();
}

void scft::relax() {
    using clk = std::chrono::steady_clock;
    const std::string s = _cfg.base_dir() + (_cfg.data_file().empty() ? std::string("scft_out") : _cfg.data_file());
    const auto t0 = clk::now();
    auto tb = t0;
    int i_save = -1;
    bool is_break = false;
    const int max_iter = (int)_cfg.max_iter();
    for (int i = 1; i <= max_iter; ++i) {
        _iter = i;
        _pmodel->update();
        if (_cfg.is_display() && i % (int)_cfg.display_interval() == 0) {
            const double t = std::chrono::duration<double>(clk::now() - tb).count() / _cfg.display_interval();
            display(t);
            tb = clk::now();
        }
        if (_cfg.is_save_data()) {
            if (i % (int)_cfg.record_interval() == 0) {
                ++i_save;
                _t.push_back(_iter);
                _time.push_back(std::chrono::duration<double>(clk::now() - t0).count());
                _H.push_back(_pmodel->H());
                _residual_error.push_back(_pmodel->residual_error());
                _incomp.push_back(0.0);
            }
            if (i_save >= 0) is_break = save_data(s, i_save);
        }
        if (is_break) break;
    }
}

bool scft::save_data(const std::string s, const int i_save) {
    std::stringstream ss;
    ss << s << "_a" << _curr_a << "_b" << _curr_b << "_c" << _curr_c << "_" << _iter << ".mat";
    if (!isFiniteNumber(_H[i_save])) {  // diverged, src/scft.cc:536-543
        std::cout << std::endl << "Error: diverge" << std::endl << "Iter: " << _iter << std::endl
                  << "H = " << _H[i_save] << std::endl;
        return true;
    }
    const bool converged = _iter > (int)_cfg.min_iter() && i_save > 0 && _residual_error[i_save] < _cfg.thresh_residual();
    if (converged || _iter == (int)_cfg.max_iter() || _iter % (int)_cfg.save_interval() == 0) {
        _pmodel->save(ss.str());
        save_record(ss.str() + ".record.mat");
        return converged || _iter == (int)_cfg.max_iter();
    }
    return false;
}

void scft::save_record(const std::string file) {
    mat5::Writer w(file);
    if (!w.ok()) return;
    const std::vector<int> dims = {(int)_t.size(), 1};
    w.put("t", dims, _t.data());
    w.put("time", dims, _time.data());
    w.put("F", dims, _H.data());
    w.put("err_residual", dims, _residual_error.data());
    w.put("incomp", dims, _incomp.data());
}
