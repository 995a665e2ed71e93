// Internal declarations shared by the host translation units.
#pragma once

#include <stdint.h>

#include "pierce_hip.h"

namespace pz {
// Learned output-size prediction for a join (mlp.cpp).
uint64_t mlp_predict_count(pz_ctx* ctx, const pz_mlp* model, const pz_dataset* a, const pz_dataset* b,
                           int predicate);
// Sum over the parts of a batch of the predicted candidate-pair counts.
uint64_t mlp_predict_batch_total(pz_ctx* ctx, const pz_mlp* model, const pz_dataset* a, const pz_dataset* b);
uint8_t pair_meta_of(const pz_join_desc& d);
}  // namespace pz
