// UnitCell.h -- cell description handed to the GPU library.
// Same class name and Config constructor as the reference (include/UnitCell.h:44-82);
// the shape and reciprocal matrices themselves are built by the library
// (the restatement of src/UnitCell.cc:142-296 lives next to the kernels).
#pragma once
#include "Config.h"
#include "polystring_amd.h"

class UnitCell {
public:
    UnitCell() {}
    explicit UnitCell(const Config &cfg)
        : _dim((int)cfg.dim()), _cs(cfg.get_crystal_system_type()), _type(cfg.get_crystal_system_type_string()),
          _a(cfg.a()), _b(cfg.b()), _c(cfg.c()), _alpha(cfg.alpha()), _beta(cfg.beta()), _gamma(cfg.gamma()) {
        // lx, ly, lz as UnitCell::calc_shape_matrix{1,2,3} set _l (src/UnitCell.cc:142-296)
        _l[0] = _a;
        _l[1] = _dim >= 2 ? (_cs == CrystalSystemType::SQUARE || _cs == CrystalSystemType::HEXAGONAL ||
                                     _cs == CrystalSystemType::CUBIC || _cs == CrystalSystemType::TETRAGONAL ||
                                     _cs == CrystalSystemType::TRIGONAL
                                 ? _a
                                 : _b)
                           : 0.0;
        _l[2] = _dim == 3 ? (_cs == CrystalSystemType::CUBIC || _cs == CrystalSystemType::TRIGONAL ? _a : _c) : 0.0;
    }
    int dim() const { return _dim; }
    std::string type() const { return _type; }
    double a() const { return _a; }
    double lx() const { return _l[0]; }
    double ly() const { return _l[1]; }
    double lz() const { return _l[2]; }

    po_grid_desc grid_desc(int Lx, int Ly, int Lz) const {
        po_grid_desc g;
        g.dim = _dim;
        g.Lx = Lx;
        g.Ly = Ly;
        g.Lz = Lz;
        g.crystal = (int)_cs;
        g.a = _a;
        g.b = _b;
        g.c = _c;
        g.alpha = _alpha;
        g.beta = _beta;
        g.gamma = _gamma;
        return g;
    }

private:
    int _dim = 3;
    CrystalSystemType _cs = CrystalSystemType::CUBIC;
    std::string _type;
    double _a = 0, _b = 0, _c = 0, _alpha = 0, _beta = 0, _gamma = 0;
    double _l[3] = {0, 0, 0};
};
