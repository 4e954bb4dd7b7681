// Model_AB.h -- AB diblock melt, Anderson mixing, periodic cell
// (include/Model_AB.h:48-114), run by the fused GPU iteration of the C ABI.
#pragma once
#include <string>
#include <vector>

#include "Model.h"
#include "UnitCell.h"
#include "polystring_amd.h"

class Model_AB : public Model {
public:
    explicit Model_AB(const std::string config_file);
    explicit Model_AB(const Config &cfg);   // used by scft/AB_ps_pbc.cc:46
    ~Model_AB();
    Model_AB(const Model_AB &) = delete;
    Model_AB &operator=(const Model_AB &) = delete;

    void init() override;
    void update() override;                     // src/Model_AB.cc:136-261
    double Hw() const override;                 // src/Model_AB.cc:263-292
    double Hs() const override;                 // src/Model_AB.cc:294-296
    double H() const override;                  // src/Model_AB.cc:298-300
    double residual_error() const override;     // src/Model_AB.cc:307-315
    double incomp() const override;             // src/Model_AB.cc:302-305
    void display() const override;              // src/Model_AB.cc:383-427
    void display_parameters() const override;
    void save(const std::string file) override;
    void save_model(const std::string file) override;
    void save_field(const std::string file) override;
    void save_density(const std::string file) override;
    void save_q(const std::string file) override;
    bool set_cell(double a, double b, double c) override;

    double Q() const;
    std::vector<double> field(int which) const;  // PO_FIELD_WA / WB / PHIA / PHIB
    po_model *handle() const { return _m; }
    // the device state changed outside update() (a batched update): the
    // cached observables are re-read on the next query
    void invalidate() { _obs_valid = false; }

private:
    void observe() const;
    po_model *_m = nullptr;
    UnitCell _uc;
    int _Lx = 1, _Ly = 1, _Lz = 1;
    double _fA = 0, _fB = 0, _chiN = 0, _dsA = 0, _dsB = 0;
    unsigned _sA = 0, _sB = 0;
    mutable double _obs[5] = {0, 0, 0, 0, 0};
    mutable bool _obs_valid = false;
};
