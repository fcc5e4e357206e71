// C-ABI host side of the spin-moment reduction (include/misamplers.h): the
// backward pass of fixed-beta sampling.  No reference counterpart: the
// reference samples but does not differentiate.  The plan fixes the model
// sizes, gives every block one tile of 64 reads and owns the int8 per-tile
// partial-sum buffer, so runs allocate nothing and are asynchronous on the
// caller's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <functional>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "host.h"
#include "misamplers.h"
#include "moments_engine.h"
#include "sa_engine.h"

using namespace misa_host;

namespace {

// LDS of one block: the staged tile (64 rows of ceil(n/4) + 1 dwords), then
// n 64-bit masks (8-byte aligned).
size_t moment_lds(int n) { return 4 * (size_t)(((n + 3) / 4) + 1) * 64 + 16 + 8 * (size_t)n; }

}  // namespace

struct misa_moments_plan {
    std::vector<misa::MomentModel> models;
    std::vector<misa::MomentBlock> blocks;
    int64_t total_vars = 0, total_edges = 0, total_samples = 0, total_part = 0, total_sums = 0;
    int max_w = 0, lds = 0;
    // blocks sorted by LDS need; launch group g covers blocks
    // [group_start[g], group_start[g + 1]) with group_lds[g] bytes each
    std::vector<int> group_start, group_lds;
    misa::MomentModel* d_models = nullptr;
    misa::MomentBlock* d_blocks = nullptr;
    int8_t* d_part = nullptr;
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
    int64_t total_tiles = 0;
    for (int b = 0; b < num_models; ++b) {
        if (num_vars[b] < 0 || num_edges[b] < 0 || num_reads[b] < 0 || num_edges[b] > INT32_MAX / 2) {
            delete pl;
            return fail(MISA_ERR_ARG, "negative model size");
        }
        const int64_t w = num_vars[b] + num_edges[b];
        const size_t lds = moment_lds(num_vars[b]);
        if (lds > lmax || w > INT32_MAX / 64) {
            delete pl;
            return fail(MISA_ERR_UNSUPPORTED, "model " + std::to_string(b) + " too large for the moment kernel");
        }
        total_tiles += (num_reads[b] + 63) / 64;
    }
    for (int b = 0; b < num_models; ++b) {
        misa::MomentModel md{};
        md.n = num_vars[b];
        md.m = (int)num_edges[b];
        md.R = num_reads[b];
        const int w = md.n + md.m;
        const int tiles = (md.R + 63) / 64;
        const int splits = std::max(tiles, 1);
        md.splits = splits;
        md.sample_off = pl->total_samples;
        md.edge_off = pl->total_edges;
        md.h_off = pl->total_vars;
        md.part_off = pl->total_part;
        md.sum_off = pl->total_sums;
        for (int s = 0; s < tiles; ++s) {
            misa::MomentBlock bk{};
            bk.model = b;
            bk.tile_lo = s;
            bk.tile_hi = s + 1;
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
        pl->lds = std::max(pl->lds, (int)moment_lds(md.n));
        pl->models.push_back(md);
    }
    // Small models must not pay the LDS of the largest: blocks are grouped by
    // LDS need (powers of two from 4 KB), one launch per group.
    {
        auto cls = [&](const misa::MomentBlock& bk) {
            const size_t need = moment_lds(pl->models[bk.model].n);
            int c = 0;
            while (((size_t)4096 << c) < need) ++c;
            return c;
        };
        std::stable_sort(pl->blocks.begin(), pl->blocks.end(),
                         [&](const misa::MomentBlock& a, const misa::MomentBlock& b) { return cls(a) < cls(b); });
        for (size_t k = 0; k < pl->blocks.size(); ++k) {
            const int c = cls(pl->blocks[k]);
            if (k == 0 || c != cls(pl->blocks[k - 1])) {
                pl->group_start.push_back((int)k);
                pl->group_lds.push_back(0);
            }
            pl->group_lds.back() = std::max(pl->group_lds.back(), (int)moment_lds(pl->models[pl->blocks[k].model].n));
        }
        pl->group_start.push_back((int)pl->blocks.size());
    }
    int rc;
    std::vector<int8_t> zero_part(std::max<int64_t>(pl->total_part, 1), 0);
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
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    for (size_t g = 0; g + 1 < pl->group_start.size(); ++g)
        MISA_HIP(misa::launch_moments_partial(pl->d_blocks + pl->group_start[g],
                                              pl->group_start[g + 1] - pl->group_start[g], pl->d_models,
                                              pl->group_lds[g], pl->group_lds[g] > 16384 ? 512 : 256, d_samples,
                                              d_edges, pl->d_part, st));
    MISA_HIP(misa::launch_moments_reduce(pl->d_models, (int)pl->models.size(), pl->max_w, pl->d_part, d_scale,
                                         d_grad_h, d_grad_J, d_sums, st));
    return MISA_OK;
}

}  // extern "C"
