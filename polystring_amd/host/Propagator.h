// Propagator.h -- Grid, Updater, RQM4, PseudoSpectral, Propagator, Density,
// Simpson and Quad4_Closed over the GPU library.
//
// Class names, constructors and method names follow the reference
// (include/Grid.h:61-183, include/Updater.h:44-69, include/RQM4.h:40-60,
// include/PseudoSpectral.h:48-70, include/Propagator.h:47-89,
// include/Density.h:51-88, include/Simpson.h, include/Quadrature4.h).
// Arrays cross as C-order double buffers (the reference's blitz layout)
// instead of blitz::Array objects; contour slices stay on the GPU and are
// copied to the host only when asked for.
#pragma once
#include <cmath>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "Config.h"
#include "UnitCell.h"
#include "polystring_amd.h"

// process-wide GPU context (device and precision from the first Config seen)
po_ctx *po_host_context(const Config *cfg = nullptr);
void po_host_check(int rc, const char *what);

// A field on the host grid (Grid::quadrature = mean on regular grids,
// src/Grid.cc:329-338).
class Grid {
public:
    Grid() {}
    Grid(const std::string name, const Config &cfg)
        : _name(name), _uc(cfg), _Lx(cfg.Lx()), _Ly(cfg.Ly()), _Lz(cfg.Lz()), _data((size_t)_Lx * _Ly * _Lz, 0.0) {}
    Grid(const std::string name, const Config &cfg, double val) : Grid(name, cfg) { _data.assign(_data.size(), val); }
    Grid(const UnitCell &uc, int Lx, int Ly, int Lz, double val)
        : _uc(uc), _Lx(Lx), _Ly(Ly), _Lz(Lz), _data((size_t)Lx * Ly * Lz, val) {}
    int Lx() const { return _Lx; }
    int Ly() const { return _Ly; }
    int Lz() const { return _Lz; }
    size_t size() const { return _data.size(); }
    const UnitCell &uc() const { return _uc; }
    const std::string &name() const { return _name; }
    double *data() { return _data.data(); }
    const double *data() const { return _data.data(); }
    double &operator()(size_t i) { return _data[i]; }
    double quadrature() const {
        double s = 0.0;
        for (double v : _data) s += v;
        return s / (double)_data.size();
    }

protected:
    std::string _name;
    UnitCell _uc;
    int _Lx = 1, _Ly = 1, _Lz = 1;
    std::vector<double> _data;
};

class Propagator;

// Updater plug-in interface (include/Updater.h:44-69).  The propagator
// updaters carry the MDE algorithm and ds; the quadrature updaters the rule.
class Updater {
public:
    virtual ~Updater() {}
    virtual void solve(Propagator &q, const Grid &w);
    virtual void solve(double *data, const Propagator &q, const Propagator &qc, const double cc) const;
    virtual Updater *clone() const = 0;
    virtual int mde_algo() const { return -1; }
    virtual int quad_rule() const { return -1; }
};

class RQM4 : public Updater {
public:
    RQM4() {}
    RQM4(const UnitCell &uc, const int Lx, const int Ly, const int Lz, const double ds) : _uc(uc), _ds(ds) {
        (void)Lx; (void)Ly; (void)Lz;
    }
    RQM4 *clone() const override { return new RQM4(*this); }
    int mde_algo() const override { return PO_RQM4; }

private:
    UnitCell _uc;
    double _ds = 0.0;
};

class PseudoSpectral : public Updater {
public:
    PseudoSpectral() {}
    PseudoSpectral(const UnitCell &uc, const int Lx, const int Ly, const int Lz, const double ds) : _uc(uc), _ds(ds) {
        (void)Lx; (void)Ly; (void)Lz;
    }
    PseudoSpectral *clone() const override { return new PseudoSpectral(*this); }
    int mde_algo() const override { return PO_OS2; }

private:
    UnitCell _uc;
    double _ds = 0.0;
};

// Etdrk4_PBC(uc, Lx, Ly, Lz, ds, M = 32) (include/Etdrk4_PBC.h:47-50): the
// Cox-Matthews ETDRK4 updater.  Only M = 32 contour points are built.
class Etdrk4_PBC : public Updater {
public:
    Etdrk4_PBC() {}
    Etdrk4_PBC(const UnitCell &uc, const int Lx, const int Ly, const int Lz, const double ds, const int M = 32)
        : _uc(uc), _ds(ds), _M(M) {
        (void)Lx; (void)Ly; (void)Lz;
    }
    Etdrk4_PBC *clone() const override { return new Etdrk4_PBC(*this); }
    int mde_algo() const override { return PO_ETDRK4; }
    int contour_points() const { return _M; }

private:
    UnitCell _uc;
    double _ds = 0.0;
    int _M = 32;
};

class Simpson : public Updater {
public:
    Simpson *clone() const override { return new Simpson; }
    int quad_rule() const override { return PO_QUAD_SIMPSON; }
};

class Quad4_Closed : public Updater {
public:
    Quad4_Closed *clone() const override { return new Quad4_Closed; }
    int quad_rule() const override { return PO_QUAD_QUAD4C; }
};

// Contour slices q(s), s = 0 .. len-1, on the GPU (include/Propagator.h:47-89).
// Like the reference, the propagator does not own its Updater
// (src/Propagator.cc:107-116); without one it runs Etdrk4_PBC, the
// reference's default updater.
class Propagator : public Grid {
public:
    Propagator(const std::string name, const Config &cfg, const int len, const double ds, Updater *up = nullptr);
    Propagator(const std::string name, const Config &cfg, const int len, const double ds, const Grid &head,
               Updater *up = nullptr);
    ~Propagator();
    Propagator(const Propagator &) = delete;
    Propagator &operator=(const Propagator &) = delete;

    void set_head(const Grid &u);
    void set_head(const double *u);
    void set_head(const std::vector<double> &u) { set_head(u.data()); }
    std::vector<double> get_tail() const { return (*this)[_length - 1]; }
    std::vector<double> operator[](const int s) const;
    void update(const Grid &w);                    // Propagator::update (src/Propagator.cc:77-79)
    void update(const Grid &w, const Updater *up);
    double Q(const int s) const;                   // src/Propagator.cc:85-89
    double Qt() const { return Q(_length - 1); }
    int len() const { return _length; }
    double ds() const { return _ds; }
    po_prop *handle() const { return _h; }

private:
    void ensure(int algo);
    int _length;
    double _ds;
    Updater *_updater;
    po_prop *_h = nullptr;
    int _algo = -1;
    std::vector<double> _pending_head;
};

// Density with normaliser cc (include/Density.h:51-88); clones and owns its
// quadrature Updater, trapezoid without one (src/Density.cc:73-121).
class Density : public Grid {
public:
    Density(const std::string name, const Config &cfg, Updater *up = nullptr)
        : Grid(name, cfg), _updater(up ? up->clone() : nullptr) {}
    ~Density() { delete _updater; }
    void set_cc(const double cc) { _cc = cc; }
    double cc() const { return _cc; }
    void update(const Propagator &q, const Propagator &qc);
    void update(const Propagator &q, const Propagator &qc, const double cc) {
        _cc = cc;
        update(q, qc);
    }

private:
    Updater *_updater;
    double _cc = 1.0;
};
