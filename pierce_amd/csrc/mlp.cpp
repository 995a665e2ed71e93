// Learned output sizing (placeholder until the estimator lands).
#include "pz_internal.h"

uint64_t pz::mlp_predict_count(pz_ctx*, const pz_mlp*, const pz_dataset*, const pz_dataset*, int) { return 0; }

uint64_t pz::mlp_predict_batch_total(pz_ctx*, const pz_mlp*, const pz_dataset*, const pz_dataset*) { return 0; }
