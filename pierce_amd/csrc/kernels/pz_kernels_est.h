// Launchers of the learned-sizing kernels (estimator.hip): join features from
// object boxes, bf16 MFMA inference of the selectivity MLP, fp32 training.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace pzk {

// ------------------------------------------------------------------ network
// 32 features -> 128 (ReLU) -> 128 (ReLU) -> 2 outputs:
//   y0 = log2(1 + candidate object pairs), y1 = log2(1 + result pairs).
// Flat fp32 parameter vector: W1 [128][32], b1 [128], W2 [128][128], b2 [128],
// W3 [2][128], b3 [2] (row-major, row = output unit).
constexpr int kMlpIn = 32, kMlpHid = 128, kMlpOut = 2;
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kMlpHid * kMlpIn;    // 4096
constexpr int kOffW2 = kOffB1 + kMlpHid;             // 4224
constexpr int kOffB2 = kOffW2 + kMlpHid * kMlpHid;   // 20608
constexpr int kOffW3 = kOffB2 + kMlpHid;             // 20736
constexpr int kOffB3 = kOffW3 + kMlpOut * kMlpHid;   // 20992
constexpr int kMlpParams = kOffB3 + kMlpOut;         // 20994
// bf16 copy for inference: W1 [128][32], W2 [128][128], W3 padded to [16][128].
constexpr int kW16Elems = kMlpHid * kMlpIn + kMlpHid * kMlpHid + 16 * kMlpHid;
// Input normalisation: mean[32], 1/std[32].
constexpr int kNormFloats = 2 * kMlpIn;

// ----------------------------------------------------------------- features
constexpr int kFeatGrid = 16;                                   // cells per axis of a part's grid
constexpr int kFeatCells = kFeatGrid * kFeatGrid * kFeatGrid;  // 4096
constexpr int kFeatBins = 16;
// Device scratch of join_features for nparts parts.
size_t features_scratch_bytes(uint32_t nparts);
// X[p][32] for each part p (a plain dataset is one part: part_of == nullptr).
// meta_of_part: predicate | dir << 2 per part (device) or nullptr -> meta.
void join_features(const uint32_t* oboxA, const uint32_t* triA, int32_t nA, const uint32_t* partA,
                   const uint32_t* oboxB, const uint32_t* triB, int32_t nB, const uint32_t* partB, uint32_t nparts,
                   const uint8_t* meta_of_part, uint8_t meta, void* scratch, float* X, hipStream_t s);

// Grid-cell estimate of pierce/query/src/cuda/estimated_overlap.cu:8-77 over n
// matched cells (GridCell layout, 4 words each): *total (device) = the sum of
// the per-cell terms in double, in a fixed order.
void cell_selectivity(const uint32_t* cells_a, const uint32_t* cells_b, uint32_t n, float cell_volume, float eps,
                      float gamma, double* total, hipStream_t s);

// Candidate pairs per part of mesh1 (pairs packed mesh1 << 32 | mesh2); counts zeroed here.
void part_pair_counts(const uint64_t* pairs, uint64_t np, const uint32_t* part_of_a, uint32_t nparts,
                      unsigned long long* counts, hipStream_t s);

// ---------------------------------------------------------------- inference
// fp32 parameters -> bf16 (round to nearest even) inference weights.
void mlp_pack_bf16(const float* params, uint16_t* w16, hipStream_t s);
// Y[m][2] = MLP(normalise(X[m][32])) with bf16 operands and fp32 accumulation
// on the matrix cores (16 rows per wavefront).
void mlp_infer_bf16(const float* X, uint32_t m, const float* norm, const uint16_t* w16, const float* params,
                    float* Y, hipStream_t s);

// ----------------------------------------------------------------- training
// Per-batch work area: normalised inputs, activations, deltas, per-sample loss.
size_t mlp_work_floats(uint32_t batch);
// One block per sample: forward (fp32) and, when T is given, backward to the
// per-sample deltas.  Sample b of the step reads row perm[s * batch + b] with
// s = ctr[0] % steps_per_epoch (ctr == nullptr: s = 0).  Y (optional) gets the
// outputs.  Loss = mean over the batch of sum over outputs of (y - t)^2.
void mlp_forward_backward(const float* X, const float* T, const uint32_t* perm, const uint32_t* ctr,
                          uint32_t steps_per_epoch, uint32_t batch, const float* norm, const float* params,
                          float* work, float* Y, hipStream_t s);
// Batch gradient of every parameter (fixed summation order over the batch).
// adam == false: grad[p] = gradient.  adam == true: AdamW step with
// hyper = {lr, beta1, beta2, eps, weight_decay} (device) and step t = ctr[1] + 1.
void mlp_grad(const float* work, uint32_t batch, float* params, float* grad, bool adam, float* m, float* v,
              const float* hyper, const uint32_t* ctr, hipStream_t s);
// step_loss[ctr[1]] = batch loss; ctr[0] += 1, ctr[1] += 1.
void mlp_step_end(const float* work, uint32_t batch, float* step_loss, uint32_t* ctr, hipStream_t s);

}  // namespace pzk
