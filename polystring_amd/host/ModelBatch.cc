// ModelBatch.cc -- see ModelBatch.h.
#include "ModelBatch.h"

#include "Propagator.h"

ModelBatch::ModelBatch(const std::vector<Model_AB *> &models, bool adjoint) : _ms(models) {
    std::vector<po_model *> hs;
    for (auto *m : _ms) hs.push_back(m->handle());
    po_host_check(po_batch_create(hs.data(), (int)hs.size(), adjoint ? 1 : 0, &_b), "po_batch_create");
}

ModelBatch::~ModelBatch() {
    if (_b) po_batch_destroy(_b);
}

void ModelBatch::update(int n) {
    po_host_check(po_batch_update(_b, n), "po_batch_update");
    for (auto *m : _ms) m->invalidate();
}

void ModelBatch::relax(int max_iter, int min_iter, int record_interval, double thresh, std::vector<int> *iters) {
    std::vector<int> it(_ms.size(), 0);
    po_host_check(po_batch_relax(_b, max_iter, min_iter, record_interval, thresh, it.data(), nullptr, nullptr),
                  "po_batch_relax");
    for (auto *m : _ms) m->invalidate();
    if (iters) *iters = it;
}
