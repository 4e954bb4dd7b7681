// Model.h -- the Model interface the scft driver runs (include/Model.h:46-87),
// without the string-method pure virtuals that make the fork's Model_AB
// abstract, so drivers such as scft/AB_ps_pbc.cc compile against it.
#pragma once
#include <string>

#include "Config.h"

class Model {
public:
    Model() {}
    explicit Model(const std::string config_file) : _cfg(config_file.c_str()) {}
    explicit Model(const Config &cfg) : _cfg(cfg) {}
    virtual ~Model() {}
    virtual void init() = 0;
    virtual void update() = 0;
    virtual double Hw() const = 0;
    virtual double Hs() const = 0;
    virtual double H() const = 0;
    virtual double residual_error() const = 0;
    virtual double incomp() const = 0;
    virtual void display() const = 0;
    virtual void display_parameters() const = 0;
    virtual void save(const std::string file) = 0;
    virtual void save_model(const std::string file) = 0;
    virtual void save_field(const std::string file) = 0;
    virtual void save_density(const std::string file) = 0;
    virtual void save_q(const std::string file) = 0;
    // New cell lengths with the current fields kept and the mixing history
    // restarted (scft::reset_model -> Model::reset, src/scft.cc:356-380);
    // false where the model cannot change its cell.
    virtual bool set_cell(double a, double b, double c) { (void)a; (void)b; (void)c; return false; }

protected:
    Config _cfg;
};
