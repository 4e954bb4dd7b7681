// ModelBatch.h -- several Model_AB systems of one grid and contour length
// updated together: one propagator launch program advances the chains of
// all of them (po_batch_create / po_batch_update).  No counterpart in the
// reference, which runs one system per process; each Model_AB keeps its own
// fields, history and observables and stays usable on its own.  The batch
// must go before its models (po_model_destroy refuses while it exists).
#pragma once
#include <vector>

#include "Model_AB.h"

class ModelBatch {
public:
    // adjoint: the cell stress fused into every batched update
    explicit ModelBatch(const std::vector<Model_AB *> &models, bool adjoint = false);
    ~ModelBatch();
    ModelBatch(const ModelBatch &) = delete;
    ModelBatch &operator=(const ModelBatch &) = delete;

    // n Model_AB::update() of every model (src/Model_AB.cc:136-261), queued
    // without host synchronisation
    void update(int n = 1);
    // scft::relax of every model (src/scft.cc:382-422,536-568), each stopping
    // on its own; iters (optional) gets each model's iteration count
    void relax(int max_iter, int min_iter, int record_interval, double thresh, std::vector<int> *iters = nullptr);
    std::size_t size() const { return _ms.size(); }

private:
    std::vector<Model_AB *> _ms;
    po_batch *_b = nullptr;
};
