// C-ABI host side of the spin-moment reduction (include/misamplers.h): the
// backward pass of fixed-beta sampling.  No reference counterpart: the
// reference samples but does not differentiate.  The plan fixes the model
// sizes, splits every model's reads over blocks (about four blocks per
// compute unit overall) and owns the int32 partial-sum buffer, so runs
// allocate nothing and are asynchronous on the caller's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "host.h"
#include "misamplers.h"
#include "moments_engine.h"
#include "sa_engine.h"

using namespace misa_host;

struct misa_moments_plan {
    std::vector<misa::MomentModel> models;
    std::vector<misa::MomentBlock> blocks;
    int64_t total_vars = 0, total_edges = 0, total_samples = 0, total_part = 0, total_sums = 0;
    int max_w = 0, lds = 0;
    misa::MomentModel* d_models = nullptr;
    misa::MomentBlock* d_blocks = nullptr;
    int32_t* d_part = nullptr;
    ~misa_moments_plan() {
        dev_free(d_models);
        dev_free(d_blocks);
        dev_free(d_part);
    }
};

extern "C" {

int misa_moments_plan_create(int num_models, const int* num_vars, const int64_t* num_edges, const int* num_reads,
                             misa_moments_plan** out) {
    if (!out) return fail(MISA_ERR_ARG, "null output pointer");
    *out = nullptr;
    if (num_models < 0 || (num_models > 0 && (!num_vars || !num_edges || !num_reads)))
        return fail(MISA_ERR_ARG, "bad model sizes");
    auto* pl = new misa_moments_plan();
    const size_t lmax = misa::max_dynamic_lds();
    double total_work = 0.0;
    for (int b = 0; b < num_models; ++b) {
        if (num_vars[b] < 0 || num_edges[b] < 0 || num_reads[b] < 0 || num_edges[b] > INT32_MAX / 2) {
            delete pl;
            return fail(MISA_ERR_ARG, "negative model size");
        }
        const int64_t w = num_vars[b] + num_edges[b];
        const size_t lds = 8 * (size_t)num_vars[b] + 4 * (size_t)w;
        if (lds > lmax) {
            delete pl;
            return fail(MISA_ERR_UNSUPPORTED, "model " + std::to_string(b) + " too large for the moment kernel");
        }
        total_work += std::ceil(num_reads[b] / 64.0) * (double)w;
    }
    const int target = 4 * device_cus();
    for (int b = 0; b < num_models; ++b) {
        misa::MomentModel md{};
        md.n = num_vars[b];
        md.m = (int)num_edges[b];
        md.R = num_reads[b];
        const int w = md.n + md.m;
        const int tiles = (md.R + 63) / 64;
        int splits = 1;
        if (total_work > 0 && tiles > 0)
            splits = (int)std::lround(target * (std::ceil(md.R / 64.0) * w) / total_work);
        splits = std::max(1, std::min(splits, std::max(tiles, 1)));
        md.splits = splits;
        md.sample_off = pl->total_samples;
        md.edge_off = pl->total_edges;
        md.h_off = pl->total_vars;
        md.part_off = pl->total_part;
        md.sum_off = pl->total_sums;
        for (int s = 0; s < splits && tiles > 0; ++s) {
            misa::MomentBlock bk{};
            bk.model = b;
            bk.tile_lo = (int)((int64_t)tiles * s / splits);
            bk.tile_hi = (int)((int64_t)tiles * (s + 1) / splits);
            bk.part_off = md.part_off + (int64_t)s * w;
            pl->blocks.push_back(bk);
        }
        if (tiles == 0) md.splits = 0;
        pl->total_samples += (int64_t)md.R * md.n;
        pl->total_edges += md.m;
        pl->total_vars += md.n;
        pl->total_part += (int64_t)md.splits * w;
        pl->total_sums += w;
        pl->max_w = std::max(pl->max_w, w);
        pl->lds = std::max(pl->lds, (int)(8 * (size_t)md.n + 4 * (size_t)w));
        pl->models.push_back(md);
    }
    int rc;
    std::vector<int32_t> zero_part(std::max<int64_t>(pl->total_part, 1), 0);
    if ((rc = dev_upload(&pl->d_models, pl->models)) || (rc = dev_upload(&pl->d_blocks, pl->blocks)) ||
        (rc = dev_upload(&pl->d_part, zero_part))) {
        delete pl;
        return rc;
    }
    *out = pl;
    return MISA_OK;
}

void misa_moments_plan_destroy(misa_moments_plan* plan) { delete plan; }

int misa_moments_run(misa_moments_plan* pl, const int8_t* d_samples, const int32_t* d_edges, const float* d_scale,
                     float* d_grad_h, float* d_grad_J, int64_t* d_sums, void* stream) {
    if (!pl) return fail(MISA_ERR_ARG, "null plan");
    if ((pl->total_samples > 0 && !d_samples) || (pl->total_edges > 0 && !d_edges))
        return fail(MISA_ERR_ARG, "null buffer");
    MISA_HIP(misa::launch_moments(pl->d_blocks, (int)pl->blocks.size(), pl->d_models, (int)pl->models.size(),
                                  pl->max_w, pl->lds, d_samples, d_edges, pl->d_part, d_scale, d_grad_h, d_grad_J,
                                  d_sums, reinterpret_cast<hipStream_t>(stream)));
    return MISA_OK;
}

}  // extern "C"
