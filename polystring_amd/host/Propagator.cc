// Propagator.cc -- host side of the Grid / Updater / Propagator / Density
// mirror; every solve and quadrature runs on the GPU through the C ABI.
#include "Propagator.h"

#include <cstdlib>
#include <iostream>

void po_host_check(int rc, const char *what) {
    if (rc != PO_OK) {
        std::cerr << "polystring_amd: " << what << " failed (" << rc << "): " << po_last_error() << std::endl;
        std::exit(1);
    }
}

po_ctx *po_host_context(const Config *cfg) {
    static po_ctx *ctx = nullptr;
    if (!ctx) {
        const int dev = cfg ? cfg->device() : 0;
        const int prec = cfg ? cfg->precision() : PO_FP64;
        po_host_check(po_ctx_create(dev, prec, &ctx), "po_ctx_create");
    }
    return ctx;
}

void Updater::solve(Propagator &q, const Grid &w) { q.update(w, this); }

void Updater::solve(double *data, const Propagator &q, const Propagator &qc, const double cc) const {
    po_host_check(po_density(q.handle(), qc.handle(), quad_rule(), cc, data), "density quadrature");
}

Propagator::Propagator(const std::string name, const Config &cfg, const int len, const double ds, Updater *up)
    : Grid(name, cfg), _length(len), _ds(ds), _updater(up) {
    po_host_context(&cfg);
}

Propagator::Propagator(const std::string name, const Config &cfg, const int len, const double ds, const Grid &head,
                       Updater *up)
    : Propagator(name, cfg, len, ds, up) {
    set_head(head);
}

Propagator::~Propagator() {
    if (_h) po_prop_destroy(_h);
}

void Propagator::ensure(int algo) {
    if (_h && algo == _algo) return;
    if (_h) po_prop_destroy(_h);
    const po_grid_desc g = _uc.grid_desc(_Lx, _Ly, _Lz);
    po_host_check(po_prop_create(po_host_context(), &g, algo, _length, _ds, &_h), "po_prop_create");
    _algo = algo;
    if (!_pending_head.empty()) po_host_check(po_prop_set_head(_h, _pending_head.data()), "set_head");
}

void Propagator::set_head(const Grid &u) { set_head(u.data()); }

void Propagator::set_head(const double *u) {
    _pending_head.assign(u, u + size());
    if (_h) po_host_check(po_prop_set_head(_h, u), "set_head");
}

std::vector<double> Propagator::operator[](const int s) const {
    std::vector<double> out(size());
    if (!_h) throw std::runtime_error("Propagator slice read before any solve");
    po_host_check(po_prop_get_slice(_h, s, out.data()), "get_slice");
    return out;
}

void Propagator::update(const Grid &w) { update(w, _updater); }

void Propagator::update(const Grid &w, const Updater *up) {
    const int algo = up && up->mde_algo() >= 0 ? up->mde_algo() : PO_ETDRK4;
    if (const auto *e = dynamic_cast<const Etdrk4_PBC *>(up))
        if (e->contour_points() != 32) {
            std::cerr << "Etdrk4_PBC: only M = 32 contour points are built" << std::endl;
            std::exit(1);
        }
    ensure(algo);
    po_host_check(po_prop_solve(_h, w.data()), "propagator solve");
}

double Propagator::Q(const int s) const {
    double q = 0.0;
    po_host_check(po_prop_Q(_h, s, &q), "Q");
    return q;
}

void Density::update(const Propagator &q, const Propagator &qc) {
    const int rule = _updater ? _updater->quad_rule() : PO_QUAD_TRAP;
    const int rc = po_density(q.handle(), qc.handle(), rule, _cc, _data.data());
    if (rc == PO_ERR_QUADRATURE) {  // src/Simpson.cc:19-26: message, density untouched
        std::cerr << po_last_error() << std::endl;
        return;
    }
    po_host_check(rc, "density");
}
