// Learned output sizing: the selectivity MLP object, its training loop on the
// GPU, and the per-part predictions the joins size their buffers with.
//
// The reference sizes its hash table from a closed-form grid estimate
// (pierce/query/src/applications/pierce_overlap_main.cpp:376-501) or counts
// exactly in a first pass (pierce_overlap_two_pass_main.cpp:39-156).  This
// model replaces the estimate: features of the two sets' object boxes go
// through a 32-128-128-2 MLP predicting log2(1 + candidate pairs) and
// log2(1 + result pairs); the buffer holds the prediction raised by a margin
// calibrated on the training data, and an overflow falls back to the exact
// two-pass count.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <random>

#include "kernels/pz_kernels.h"
#include "kernels/pz_kernels_est.h"
#include "pz_internal.h"

using namespace pz;

static_assert(pzk::kMlpParams == PZ_MLP_PARAMS, "parameter count");
static_assert(pzk::kMlpIn == PZ_MLP_FEATURES, "feature count");
static_assert(pzk::kMlpOut == PZ_MLP_OUTPUTS, "output count");

struct pz_mlp {
    pz_ctx* ctx = nullptr;
    int device = 0;
    float* d_params = nullptr;  // authoritative parameters (fp32)
    float* d_norm = nullptr;    // mean[32], 1/std[32]
    uint16_t* d_w16 = nullptr;  // bf16 inference weights, valid when packed
    float* d_m = nullptr;       // AdamW moments
    float* d_v = nullptr;
    bool packed = false;
    bool norm_set = false;
    float norm[pzk::kNormFloats];
    float margin[PZ_MLP_OUTPUTS] = {1.0f, 1.0f};  // log2 margins of the sizing upper bounds
};

namespace {

constexpr char kMagic[8] = {'P', 'Z', 'M', 'L', 'P', '0', '0', '1'};

void upload(pz_ctx* c, void* dst, const void* src, size_t bytes) {
    HIP_OK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->s));
    HIP_OK(hipStreamSynchronize(c->s));
}
void download(pz_ctx* c, void* dst, const void* src, size_t bytes) {
    HIP_OK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->s));
    HIP_OK(hipStreamSynchronize(c->s));
}

void ensure_packed(pz_mlp* m) {
    if (m->packed) return;
    pzk::mlp_pack_bf16(m->d_params, m->d_w16, m->ctx->s);
    m->packed = true;
}

void set_norm(pz_mlp* m, const float* norm) {
    std::memcpy(m->norm, norm, sizeof(m->norm));
    upload(m->ctx, m->d_norm, m->norm, sizeof(m->norm));
    m->norm_set = true;
}

// He-uniform hidden layers, small output layer, zero biases.
std::vector<float> init_params(uint64_t seed) {
    std::vector<float> p(pzk::kMlpParams, 0.f);
    std::mt19937_64 g(seed);
    auto fill = [&](int off, int rows, int cols, float scale) {
        std::uniform_real_distribution<float> u(-scale, scale);
        for (int i = 0; i < rows * cols; ++i) p[off + i] = u(g);
    };
    fill(pzk::kOffW1, pzk::kMlpHid, pzk::kMlpIn, std::sqrt(6.f / pzk::kMlpIn));
    fill(pzk::kOffW2, pzk::kMlpHid, pzk::kMlpHid, std::sqrt(6.f / pzk::kMlpHid));
    fill(pzk::kOffW3, pzk::kMlpOut, pzk::kMlpHid, 0.1f * std::sqrt(6.f / pzk::kMlpHid));
    return p;
}

pz_mlp* new_model(pz_ctx* c) {
    std::unique_ptr<pz_mlp> m(new pz_mlp());
    m->ctx = c;
    m->device = c->device;
    HIP_OK(hipSetDevice(c->device));
    HIP_OK(hipMalloc(&m->d_params, sizeof(float) * pzk::kMlpParams));
    HIP_OK(hipMalloc(&m->d_norm, sizeof(float) * pzk::kNormFloats));
    HIP_OK(hipMalloc(&m->d_w16, sizeof(uint16_t) * pzk::kW16Elems));
    HIP_OK(hipMalloc(&m->d_m, sizeof(float) * pzk::kMlpParams));
    HIP_OK(hipMalloc(&m->d_v, sizeof(float) * pzk::kMlpParams));
    for (int k = 0; k < pzk::kMlpIn; ++k) {
        m->norm[k] = 0.f;
        m->norm[pzk::kMlpIn + k] = 1.f;
    }
    upload(c, m->d_norm, m->norm, sizeof(m->norm));
    return m.release();
}

// Features of every part of a join (device X, nparts x 32).
float* features_device(pz_ctx* c, const pz_dataset* a, const pz_dataset* b, const SetPrep& pa, const SetPrep& pb,
                       const uint8_t* meta_of_part, uint8_t meta, uint32_t* nparts_out, hipStream_t stream) {
    const bool parts = !a->part_obj_off.empty();
    if (parts != !b->part_obj_off.empty() || (parts && a->part_obj_off.size() != b->part_obj_off.size()))
        throw PzError(PZ_E_INVALID, "both datasets must be plain or batched with the same number of parts");
    const uint32_t nparts = parts ? (uint32_t)(a->part_obj_off.size() - 1) : 1u;
    void* scratch = buf<char>(c, "est.scratch", pzk::features_scratch_bytes(nparts));
    float* X = buf<float>(c, "est.X", (size_t)nparts * pzk::kMlpIn);
    pzk::join_features(pa.obox, pa.tri_cnt, a->nobj, parts ? a->part_of : nullptr, pb.obox, pb.tri_cnt, b->nobj,
                       parts ? b->part_of : nullptr, nparts, meta_of_part, meta, scratch, X, stream);
    *nparts_out = nparts;
    return X;
}

void check_model(pz_ctx* c, const pz_mlp* m) {
    if (!m) throw PzError(PZ_E_INVALID, "learned sizing needs a model");
    if (m->ctx != c) throw PzError(PZ_E_INVALID, "the model belongs to another context");
}

uint8_t meta_of_predicate(int32_t pred) {
    pz_join_desc d;
    pz_join_desc_defaults(&d);
    d.predicate = pred;
    return pair_meta_of(d);
}

// Quantile q of v (linear interpolation between order statistics).
float quantile(std::vector<float> v, float q) {
    if (v.empty()) return 0.f;
    std::sort(v.begin(), v.end());
    const double pos = q * (double)(v.size() - 1);
    const size_t i = (size_t)pos;
    const double f = pos - (double)i;
    return (float)(i + 1 < v.size() ? v[i] * (1 - f) + v[i + 1] * f : v[i]);
}

}  // namespace

uint32_t pz::mlp_predict_async(pz_ctx* c, const pz_mlp* model, const pz_dataset* a, const pz_dataset* b,
                               const SetPrep& pa, const SetPrep& pb, const uint8_t* meta_of_part, uint8_t meta) {
    check_model(c, model);
    pz_mlp* m = const_cast<pz_mlp*>(model);
    HIP_OK(hipEventRecord(c->ev_fork, c->s));
    HIP_OK(hipStreamWaitEvent(c->s2, c->ev_fork, 0));
    uint32_t np = 0;
    float* X = features_device(c, a, b, pa, pb, meta_of_part, meta, &np, c->s2);
    float* Y = buf<float>(c, "est.Y", (size_t)np * pzk::kMlpOut);
    if (!m->packed) {
        pzk::mlp_pack_bf16(m->d_params, m->d_w16, c->s2);
        m->packed = true;
    }
    pzk::mlp_infer_bf16(X, np, m->d_norm, m->d_w16, m->d_params, Y, c->s2);
    const size_t bytes = sizeof(float) * pzk::kMlpOut * np;
    if (c->pinned_est_n < bytes) {
        if (c->pinned_est) HIP_OK(hipHostFree(c->pinned_est));
        c->pinned_est = nullptr;
        HIP_OK(hipHostMalloc(&c->pinned_est, bytes, hipHostMallocDefault));
        c->pinned_est_n = bytes;
    }
    HIP_OK(hipMemcpyAsync(c->pinned_est, Y, bytes, hipMemcpyDeviceToHost, c->s2));
    HIP_OK(hipEventRecord(c->ev_join, c->s2));
    return np;
}

double pz::candidate_estimate(pz_ctx* c, const pz_dataset* a, const pz_dataset* b, const SetPrep& pa,
                              const SetPrep& pb) {
    uint32_t np = 0;
    const float* X = features_device(c, a, b, pa, pb, nullptr, 0, &np, c->s);
    const float f = readback<float>(c, {X + 12})[0];  // log2(1 + expected box pairs)
    return std::exp2(std::min(std::max((double)f, 0.0), 60.0)) - 1.0;
}

std::vector<SizePrediction> pz::mlp_predict_wait(pz_ctx* c, const pz_mlp* model, uint32_t np) {
    HIP_OK(hipEventSynchronize(c->ev_join));
    const float* h = static_cast<const float*>(c->pinned_est);
    std::vector<SizePrediction> out(np);
    for (uint32_t p = 0; p < np; ++p) {
        auto val = [](float y) { return std::exp2(std::min(std::max((double)y, 0.0), 50.0)) - 1.0; };
        const float y0 = h[2 * p], y1 = h[2 * p + 1];
        out[p] = {val(y0), val(y1), val(y0 + model->margin[0]), val(y1 + model->margin[1])};
    }
    return out;
}

extern "C" {

void pz_mlp_train_desc_defaults(pz_mlp_train_desc* d) {
    std::memset(d, 0, sizeof(*d));
    d->epochs = 150;
    d->batch = 64;
    d->lr = 2e-3f;
    d->lr_final = 1e-4f;
    d->weight_decay = 0.f;
    d->seed = 1;
    d->use_graph = 1;
    d->margin_quantile = 0.95f;
    d->validation_fraction = 0.2f;
}

pz_status pz_mlp_create(pz_ctx* c, uint64_t seed, pz_mlp** out) {
    if (!c || !out) return fail(PZ_E_INVALID, "pz_mlp_create: null argument");
    return guarded([&]() -> pz_status {
        std::unique_ptr<pz_mlp> m(new_model(c));
        std::vector<float> p = init_params(seed);
        upload(c, m->d_params, p.data(), sizeof(float) * p.size());
        *out = m.release();
        return PZ_OK;
    });
}

void pz_mlp_destroy(pz_mlp* m) {
    if (!m) return;
    (void)hipSetDevice(m->device);  // the context may already be gone
    for (void* p : {(void*)m->d_params, (void*)m->d_norm, (void*)m->d_w16, (void*)m->d_m, (void*)m->d_v})
        if (p) (void)hipFree(p);
    delete m;
}

pz_status pz_mlp_get_params(const pz_mlp* m, float* params, float* norm, float* margin) {
    if (!m) return fail(PZ_E_INVALID, "pz_mlp_get_params: null model");
    return guarded([&]() -> pz_status {
        HIP_OK(hipSetDevice(m->ctx->device));
        if (params) download(m->ctx, params, m->d_params, sizeof(float) * pzk::kMlpParams);
        if (norm) std::memcpy(norm, m->norm, sizeof(m->norm));
        if (margin) std::memcpy(margin, m->margin, sizeof(m->margin));
        return PZ_OK;
    });
}

pz_status pz_mlp_set_params(pz_mlp* m, const float* params, const float* norm, const float* margin) {
    if (!m) return fail(PZ_E_INVALID, "pz_mlp_set_params: null model");
    return guarded([&]() -> pz_status {
        HIP_OK(hipSetDevice(m->ctx->device));
        if (params) {
            upload(m->ctx, m->d_params, params, sizeof(float) * pzk::kMlpParams);
            m->packed = false;
        }
        if (norm) {
            for (int k = 0; k < pzk::kNormFloats; ++k)
                if (!std::isfinite(norm[k])) throw PzError(PZ_E_INVALID, "normalisation must be finite");
            set_norm(m, norm);
        }
        if (margin) std::memcpy(m->margin, margin, sizeof(m->margin));
        return PZ_OK;
    });
}

pz_status pz_mlp_save(const pz_mlp* m, const char* path) {
    if (!m || !path) return fail(PZ_E_INVALID, "pz_mlp_save: null argument");
    return guarded([&]() -> pz_status {
        std::vector<float> p(pzk::kMlpParams);
        HIP_OK(hipSetDevice(m->ctx->device));
        download(m->ctx, p.data(), m->d_params, sizeof(float) * p.size());
        FILE* f = std::fopen(path, "wb");
        if (!f) throw PzError(PZ_E_IO, std::string("cannot write ") + path);
        const uint32_t dims[5] = {(uint32_t)pzk::kMlpIn, (uint32_t)pzk::kMlpHid, (uint32_t)pzk::kMlpOut,
                                  (uint32_t)pzk::kMlpParams, m->norm_set ? 1u : 0u};
        bool ok = std::fwrite(kMagic, 1, 8, f) == 8 && std::fwrite(dims, 4, 5, f) == 5 &&
                  std::fwrite(m->norm, 4, pzk::kNormFloats, f) == (size_t)pzk::kNormFloats &&
                  std::fwrite(m->margin, 4, PZ_MLP_OUTPUTS, f) == (size_t)PZ_MLP_OUTPUTS &&
                  std::fwrite(p.data(), 4, p.size(), f) == p.size();
        ok = (std::fclose(f) == 0) && ok;
        if (!ok) throw PzError(PZ_E_IO, std::string("write failed: ") + path);
        return PZ_OK;
    });
}

pz_status pz_mlp_load(pz_ctx* c, const char* path, pz_mlp** out) {
    if (!c || !path || !out) return fail(PZ_E_INVALID, "pz_mlp_load: null argument");
    return guarded([&]() -> pz_status {
        FILE* f = std::fopen(path, "rb");
        if (!f) throw PzError(PZ_E_IO, std::string("cannot open ") + path);
        char magic[8];
        uint32_t dims[5];
        float norm[pzk::kNormFloats], margin[PZ_MLP_OUTPUTS];
        std::vector<float> p(pzk::kMlpParams);
        bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, kMagic, 8) == 0 &&
                  std::fread(dims, 4, 5, f) == 5 && dims[0] == (uint32_t)pzk::kMlpIn &&
                  dims[1] == (uint32_t)pzk::kMlpHid && dims[2] == (uint32_t)pzk::kMlpOut &&
                  dims[3] == (uint32_t)pzk::kMlpParams &&
                  std::fread(norm, 4, pzk::kNormFloats, f) == (size_t)pzk::kNormFloats &&
                  std::fread(margin, 4, PZ_MLP_OUTPUTS, f) == (size_t)PZ_MLP_OUTPUTS &&
                  std::fread(p.data(), 4, p.size(), f) == p.size();
        std::fclose(f);
        if (!ok) throw PzError(PZ_E_IO, std::string("not a selectivity model file: ") + path);
        std::unique_ptr<pz_mlp> m(new_model(c));
        upload(c, m->d_params, p.data(), sizeof(float) * p.size());
        set_norm(m.get(), norm);
        m->norm_set = dims[4] != 0;
        std::memcpy(m->margin, margin, sizeof(margin));
        *out = m.release();
        return PZ_OK;
    });
}

pz_status pz_join_features(pz_ctx* c, const pz_dataset* a, const pz_dataset* b, int32_t predicate, float* out) {
    if (!c || !a || !b || !out) return fail(PZ_E_INVALID, "pz_join_features: null argument");
    if (predicate < 0 || predicate > 2) return fail(PZ_E_INVALID, "unknown predicate");
    if (!a->part_obj_off.empty() || !b->part_obj_off.empty())
        return fail(PZ_E_INVALID, "pz_join_features takes plain datasets (use pz_batch_features for batches)");
    return guarded([&]() -> pz_status {
        HIP_OK(hipSetDevice(c->device));
        SetPrep pa = prep_set(c, a, "A"), pb = prep_set(c, b, "B");
        uint32_t np = 0;
        float* X = features_device(c, a, b, pa, pb, nullptr, meta_of_predicate(predicate), &np, c->s);
        download(c, out, X, sizeof(float) * pzk::kMlpIn);
        return PZ_OK;
    });
}

pz_status pz_batch_features(pz_ctx* c, const pz_dataset* a, const pz_dataset* b, uint32_t nq,
                            const int32_t* predicates, float* out) {
    if (!c || !a || !b || !predicates || !out || nq == 0) return fail(PZ_E_INVALID, "pz_batch_features: bad argument");
    if (a->part_obj_off.size() != (size_t)nq + 1 || b->part_obj_off.size() != (size_t)nq + 1)
        return fail(PZ_E_INVALID, "pz_batch_features: both datasets must be batches of nq parts");
    for (uint32_t q = 0; q < nq; ++q)
        if (predicates[q] < 0 || predicates[q] > 2) return fail(PZ_E_INVALID, "unknown predicate");
    return guarded([&]() -> pz_status {
        HIP_OK(hipSetDevice(c->device));
        std::vector<uint8_t> hm(nq);
        for (uint32_t q = 0; q < nq; ++q) hm[q] = meta_of_predicate(predicates[q]);
        uint8_t* dm = buf<uint8_t>(c, "est.meta", nq);
        upload(c, dm, hm.data(), nq);
        SetPrep pa = prep_set(c, a, "A"), pb = prep_set(c, b, "B");
        uint32_t np = 0;
        float* X = features_device(c, a, b, pa, pb, dm, 0, &np, c->s);
        download(c, out, X, sizeof(float) * pzk::kMlpIn * nq);
        return PZ_OK;
    });
}

static pz_status run_network(pz_ctx* c, const pz_mlp* m, const float* X, uint64_t n, float* Y, bool bf16) {
    if (!c || !m || !X || !Y) return fail(PZ_E_INVALID, "null argument");
    if (n > (1ull << 26)) return fail(PZ_E_LIMIT, "at most 2^26 rows per call");
    return guarded([&]() -> pz_status {
        check_model(c, m);
        if (n == 0) return PZ_OK;
        HIP_OK(hipSetDevice(c->device));
        float* dX = buf<float>(c, "mlp.X", n * pzk::kMlpIn);
        float* dY = buf<float>(c, "mlp.Y", n * pzk::kMlpOut);
        upload(c, dX, X, sizeof(float) * pzk::kMlpIn * n);
        if (bf16) {
            ensure_packed(const_cast<pz_mlp*>(m));
            pzk::mlp_infer_bf16(dX, (uint32_t)n, m->d_norm, m->d_w16, m->d_params, dY, c->s);
        } else {
            float* work = buf<float>(c, "mlp.work", pzk::mlp_work_floats((uint32_t)n));
            pzk::mlp_forward_backward(dX, nullptr, nullptr, nullptr, 1, (uint32_t)n, m->d_norm, m->d_params, work,
                                      dY, c->s);
        }
        download(c, Y, dY, sizeof(float) * pzk::kMlpOut * n);
        return PZ_OK;
    });
}

pz_status pz_mlp_predict(pz_ctx* c, const pz_mlp* m, const float* X, uint64_t n, float* Y) {
    return run_network(c, m, X, n, Y, true);
}

pz_status pz_mlp_forward(pz_ctx* c, const pz_mlp* m, const float* X, uint64_t n, float* Y) {
    return run_network(c, m, X, n, Y, false);
}

pz_status pz_mlp_gradients(pz_ctx* c, const pz_mlp* m, const float* X, const float* T, uint64_t n, float* loss,
                           float* grad) {
    if (!c || !m || !X || !T || !grad || n == 0) return fail(PZ_E_INVALID, "pz_mlp_gradients: bad argument");
    if (n > 65536) return fail(PZ_E_LIMIT, "at most 65536 samples per gradient");
    return guarded([&]() -> pz_status {
        check_model(c, m);
        HIP_OK(hipSetDevice(c->device));
        hipStream_t s = c->s;
        float* dX = buf<float>(c, "mlp.X", n * pzk::kMlpIn);
        float* dT = buf<float>(c, "mlp.T", n * pzk::kMlpOut);
        float* work = buf<float>(c, "mlp.work", pzk::mlp_work_floats((uint32_t)n));
        float* dG = buf<float>(c, "mlp.G", pzk::kMlpParams);
        upload(c, dX, X, sizeof(float) * pzk::kMlpIn * n);
        upload(c, dT, T, sizeof(float) * pzk::kMlpOut * n);
        pzk::mlp_forward_backward(dX, dT, nullptr, nullptr, 1, (uint32_t)n, m->d_norm, m->d_params, work, nullptr, s);
        pzk::mlp_grad(work, (uint32_t)n, m->d_params, dG, false, nullptr, nullptr, nullptr, nullptr, s);
        download(c, grad, dG, sizeof(float) * pzk::kMlpParams);
        if (loss) {
            std::vector<float> l(n);
            // per-sample squared errors sit at the end of the work area
            download(c, l.data(), work + pzk::mlp_work_floats((uint32_t)n) - n, sizeof(float) * n);
            double t = 0;
            for (float v : l) t += v;
            *loss = (float)(t / (double)n);
        }
        return PZ_OK;
    });
}

pz_status pz_mlp_train(pz_ctx* c, pz_mlp* m, const float* X0, const float* T0, uint64_t n,
                       const pz_mlp_train_desc* desc, float* epoch_loss) {
    if (!c || !m || !X0 || !T0 || n == 0) return fail(PZ_E_INVALID, "pz_mlp_train: bad argument");
    if (n > (1ull << 24)) return fail(PZ_E_LIMIT, "at most 2^24 training samples");
    pz_mlp_train_desc d;
    if (desc) d = *desc; else pz_mlp_train_desc_defaults(&d);
    if (d.epochs <= 0 || d.batch <= 0 || !(d.lr > 0) || !(d.lr_final >= 0) || d.weight_decay < 0 ||
        !(d.margin_quantile >= 0 && d.margin_quantile <= 1) ||
        !(d.validation_fraction >= 0 && d.validation_fraction < 1))
        return fail(PZ_E_INVALID, "pz_mlp_train: invalid hyper-parameters");
    return guarded([&]() -> pz_status {
        check_model(c, m);
        HIP_OK(hipSetDevice(c->device));
        hipStream_t s = c->s;
        std::mt19937_64 rng(d.seed);
        // validation split: a seeded random subset held out of training
        std::vector<uint32_t> order(n);
        for (uint64_t i = 0; i < n; ++i) order[i] = (uint32_t)i;
        std::shuffle(order.begin(), order.end(), rng);
        uint64_t nval = (uint64_t)std::floor(d.validation_fraction * (double)n);
        if (nval < 8 || n - nval < 1) nval = 0;
        std::vector<float> Xt, Tt, Xv, Tv;
        for (uint64_t j = 0; j < n; ++j) {
            const uint64_t i = order[j];
            auto& xs = j < nval ? Xv : Xt;
            auto& ts = j < nval ? Tv : Tt;
            xs.insert(xs.end(), X0 + i * pzk::kMlpIn, X0 + (i + 1) * pzk::kMlpIn);
            ts.insert(ts.end(), T0 + i * pzk::kMlpOut, T0 + (i + 1) * pzk::kMlpOut);
        }
        n -= nval;
        const float* X = Xt.data();
        const float* T = Tt.data();
        const uint32_t B = (uint32_t)std::min<uint64_t>((uint64_t)d.batch, n);
        const uint32_t steps = (uint32_t)(n / B);  // full batches per epoch (the remainder reshuffles in)
        for (uint64_t i = 0; i < (n + nval) * pzk::kMlpIn; ++i)
            if (!std::isfinite(X0[i])) throw PzError(PZ_E_INVALID, "non-finite feature");
        for (uint64_t i = 0; i < (n + nval) * pzk::kMlpOut; ++i)
            if (!std::isfinite(T0[i])) throw PzError(PZ_E_INVALID, "non-finite target");

        if (!m->norm_set) {  // first training: input normalisation and output bias from the data
            float norm[pzk::kNormFloats];
            for (int k = 0; k < pzk::kMlpIn; ++k) {
                double s1 = 0, s2 = 0;
                for (uint64_t i = 0; i < n; ++i) {
                    s1 += X[i * pzk::kMlpIn + k];
                    s2 += (double)X[i * pzk::kMlpIn + k] * X[i * pzk::kMlpIn + k];
                }
                const double mu = s1 / (double)n, var = std::max(s2 / (double)n - mu * mu, 0.0);
                norm[k] = (float)mu;
                norm[pzk::kMlpIn + k] = (float)(1.0 / std::max(std::sqrt(var), 1e-3));
            }
            set_norm(m, norm);
            float b3[pzk::kMlpOut];
            for (int o = 0; o < pzk::kMlpOut; ++o) {
                double t = 0;
                for (uint64_t i = 0; i < n; ++i) t += T[i * pzk::kMlpOut + o];
                b3[o] = (float)(t / (double)n);
            }
            upload(c, m->d_params + pzk::kOffB3, b3, sizeof(b3));
        }

        float* dX = buf<float>(c, "mlp.X", n * pzk::kMlpIn);
        float* dT = buf<float>(c, "mlp.T", n * pzk::kMlpOut);
        uint32_t* dperm = buf<uint32_t>(c, "mlp.perm", (size_t)steps * B);
        float* work = buf<float>(c, "mlp.work", pzk::mlp_work_floats(B));
        float* hyper = buf<float>(c, "mlp.hyper", 8);
        uint32_t* ctr = buf<uint32_t>(c, "mlp.ctr", 2);
        const uint64_t total_steps = (uint64_t)steps * (uint64_t)d.epochs;
        float* step_loss = buf<float>(c, "mlp.loss", total_steps);
        upload(c, dX, X, sizeof(float) * pzk::kMlpIn * n);
        upload(c, dT, T, sizeof(float) * pzk::kMlpOut * n);
        pzk::fill_u32(reinterpret_cast<uint32_t*>(m->d_m), 0u, pzk::kMlpParams, s);
        pzk::fill_u32(reinterpret_cast<uint32_t*>(m->d_v), 0u, pzk::kMlpParams, s);
        pzk::fill_u32(ctr, 0u, 2, s);

        auto launch_epoch = [&]() {
            for (uint32_t st = 0; st < steps; ++st) {
                pzk::mlp_forward_backward(dX, dT, dperm, ctr, steps, B, m->d_norm, m->d_params, work, nullptr, s);
                pzk::mlp_grad(work, B, m->d_params, nullptr, true, m->d_m, m->d_v, hyper, ctr, s);
                pzk::mlp_step_end(work, B, step_loss, ctr, s);
            }
        };
        hipGraphExec_t exec = nullptr;
        hipGraph_t graph = nullptr;
        struct GraphGuard {
            hipGraphExec_t* e;
            hipGraph_t* g;
            ~GraphGuard() {
                if (*e) (void)hipGraphExecDestroy(*e);
                if (*g) (void)hipGraphDestroy(*g);
            }
        } guard{&exec, &graph};
        if (d.use_graph) {
            HIP_OK(hipStreamSynchronize(s));
            HIP_OK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            launch_epoch();
            HIP_OK(hipStreamEndCapture(s, &graph));
            HIP_OK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        }
        order.resize(n);
        for (uint64_t i = 0; i < n; ++i) order[i] = (uint32_t)i;
        const float kPi = 3.14159265358979f;
        for (int32_t e = 0; e < d.epochs; ++e) {
            std::shuffle(order.begin(), order.end(), rng);
            // cosine decay from lr to lr_final over the epochs
            const float frac = d.epochs > 1 ? (float)e / (float)(d.epochs - 1) : 0.f;
            const float lr = d.lr_final + 0.5f * (d.lr - d.lr_final) * (1.f + std::cos(kPi * frac));
            const float h[5] = {lr, 0.9f, 0.999f, 1e-8f, d.weight_decay};
            HIP_OK(hipMemcpyAsync(dperm, order.data(), sizeof(uint32_t) * steps * B, hipMemcpyHostToDevice, s));
            HIP_OK(hipMemcpyAsync(hyper, h, sizeof(h), hipMemcpyHostToDevice, s));
            HIP_OK(hipStreamSynchronize(s));  // host arrays are reused next epoch
            if (exec)
                HIP_OK(hipGraphLaunch(exec, s));
            else
                launch_epoch();
        }
        m->packed = false;
        if (epoch_loss) {
            std::vector<float> l(total_steps);
            download(c, l.data(), step_loss, sizeof(float) * total_steps);
            for (int32_t e = 0; e < d.epochs; ++e) {
                double t = 0;
                for (uint32_t st = 0; st < steps; ++st) t += l[(uint64_t)e * steps + st];
                epoch_loss[e] = (float)(t / steps);
            }
        } else {
            HIP_OK(hipStreamSynchronize(s));
        }
        // Sizing margins: quantile of (target - bf16 prediction) over the
        // validation samples (the training samples when there are none).
        const uint64_t nc = nval ? nval : n;
        const float* Xc = nval ? Xv.data() : X;
        const float* Tc = nval ? Tv.data() : T;
        std::vector<float> Y(nc * pzk::kMlpOut);
        pz_status st = run_network(c, m, Xc, nc, Y.data(), true);
        if (st != PZ_OK) throw PzError(st, pz_last_error());
        for (int o = 0; o < pzk::kMlpOut; ++o) {
            std::vector<float> r(nc);
            for (uint64_t i = 0; i < nc; ++i) r[i] = Tc[i * pzk::kMlpOut + o] - Y[i * pzk::kMlpOut + o];
            m->margin[o] = std::max(quantile(r, d.margin_quantile), 0.f) + 0.05f;
        }
        return PZ_OK;
    });
}

}  // extern "C"
