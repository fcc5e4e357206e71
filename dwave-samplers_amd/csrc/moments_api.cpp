// C-ABI host side of the spin-moment reduction (include/misamplers.h): the
// backward pass of fixed-beta sampling.  No reference counterpart: the
// reference samples but does not differentiate.  The plan fixes the model
// sizes, gives every partial block one to four tiles of 64 reads of one
// model and owns the int8 per-tile partial-sum buffer, so runs allocate
// nothing and are asynchronous on the caller's stream.
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

namespace {

// Waves building one tile's masks: enough for about four 16-variable chunks
// each (a block holds 4 / wpt tiles).
int waves_per_tile(int n) {
    const int nch = (n + 15) / 16;
    return nch >= 16 ? 4 : nch >= 4 ? 2 : 1;
}

// Partial launches share one dynamic LDS size, so blocks up to this size go
// in one launch; larger ones are grouped by power-of-two classes.
constexpr size_t kSharedLaunchLds = 65536;

}  // namespace

struct misa_moments_plan {
    std::vector<misa::MomentModel> models;
    std::vector<misa::MomentBlock> blocks;
    std::vector<misa::MomentSpan> spans;
    int64_t total_vars = 0, total_edges = 0, total_samples = 0, total_part = 0, total_sums = 0;
    // launch group g covers blocks [group_start[g], group_start[g + 1]) with
    // group_lds[g] bytes each
    std::vector<int> group_start, group_lds;
    misa::MomentModel* d_models = nullptr;
    misa::MomentBlock* d_blocks = nullptr;
    misa::MomentSpan* d_spans = nullptr;
    int8_t* d_part = nullptr;
    ~misa_moments_plan() {
        dev_free(d_models);
        dev_free(d_blocks);
        dev_free(d_spans);
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
    // LDS of one tile: masks (8 bytes per variable, rounded up to 16 bytes)
    // and its partial row
    auto tile_lds = [](int n, int64_t m) {
        return ((8 * (size_t)n + 15) & ~(size_t)15) + (((size_t)n + (size_t)m + 15) & ~(size_t)15);
    };
    for (int b = 0; b < num_models; ++b) {
        if (num_vars[b] < 0 || num_edges[b] < 0 || num_reads[b] < 0 || num_edges[b] > INT32_MAX / 2) {
            delete pl;
            return fail(MISA_ERR_ARG, "negative model size");
        }
        const int64_t w = num_vars[b] + num_edges[b];
        if (std::max(tile_lds(num_vars[b], num_edges[b]), ((8 * (size_t)num_vars[b] + 15) & ~(size_t)15) +
                                                                misa::moment_stage_lds(num_vars[b])) > lmax ||
            w > INT32_MAX / 64) {
            delete pl;
            return fail(MISA_ERR_UNSUPPORTED, "model " + std::to_string(b) + " too large for the moment kernel");
        }
    }
    // LDS of a block: its tiles' masks and partial rows (pw bytes each)
    auto block_lds = [&](const misa::MomentBlock& bk) {
        const auto& md = pl->models[bk.model];
        const size_t t = (size_t)(bk.tile_hi - bk.tile_lo);
        return ((8 * (size_t)md.n * t + 15) & ~(size_t)15) + std::max(t * (size_t)md.pw, misa::moment_stage_lds(md.n)) +
               misa::moment_edge_lds((int)t, md.m);
    };
    for (int b = 0; b < num_models; ++b) {
        misa::MomentModel md{};
        md.n = num_vars[b];
        md.m = (int)num_edges[b];
        md.R = num_reads[b];
        const int w = md.n + md.m;
        md.pw = (w + 15) & ~15;
        const int tiles = (md.R + 63) / 64;
        md.splits = tiles;
        md.sample_off = pl->total_samples;
        md.edge_off = pl->total_edges;
        md.h_off = pl->total_vars;
        md.part_off = pl->total_part;
        md.sum_off = pl->total_sums;
        pl->models.push_back(md);
        int wpt = waves_per_tile(md.n);
        while (wpt < 4 && tile_lds(md.n, md.m) * (size_t)(4 / wpt) + misa::moment_stage_lds(md.n) > lmax) wpt *= 2;  // bound, not exact
        int tpb = 4 / wpt;
        if (md.n <= 16 && tile_lds(md.n, md.m) * misa::kMaxTilesPerBlock +
                                  misa::moment_edge_lds(misa::kMaxTilesPerBlock, md.m) <= lmax)
            tpb = misa::kMaxTilesPerBlock;
        for (int s = 0; s < tiles; s += tpb) {
            misa::MomentBlock bk{};
            bk.md = md;
            bk.model = b;
            bk.tile_lo = s;
            bk.tile_hi = std::min(tiles, s + tpb);
            bk.wpt = wpt;
            bk.part_off = md.part_off + (int64_t)s * md.pw;
            pl->blocks.push_back(bk);
        }
        for (int e0 = 0; e0 < w; e0 += misa::kReduceSpan) pl->spans.push_back(misa::MomentSpan{md, b, e0});
        pl->total_samples += (int64_t)md.R * md.n;
        pl->total_edges += md.m;
        pl->total_vars += md.n;
        pl->total_part += (int64_t)md.splits * md.pw;
        pl->total_sums += w;
    }
    // One launch for every block of up to kSharedLaunchLds bytes, so small
    // models run beside large ones; larger blocks by power-of-two class.
    // Heaviest blocks first inside a launch.
    {
        auto cls = [&](const misa::MomentBlock& bk) {
            const size_t need = block_lds(bk);
            int c = 0;
            while ((kSharedLaunchLds << c) < need) ++c;
            return c;
        };
        auto work = [&](const misa::MomentBlock& bk) {
            const auto& md = pl->models[bk.model];
            return (int64_t)(bk.tile_hi - bk.tile_lo) * (md.n + md.m);
        };
        std::stable_sort(pl->blocks.begin(), pl->blocks.end(), [&](const misa::MomentBlock& a, const misa::MomentBlock& b) {
            const int ca = cls(a), cb = cls(b);
            return ca != cb ? ca < cb : work(a) > work(b);
        });
        for (size_t k = 0; k < pl->blocks.size(); ++k) {
            const int c = cls(pl->blocks[k]);
            if (k == 0 || c != cls(pl->blocks[k - 1])) {
                pl->group_start.push_back((int)k);
                pl->group_lds.push_back(0);
            }
            pl->group_lds.back() = std::max(pl->group_lds.back(), (int)block_lds(pl->blocks[k]));
        }
        pl->group_start.push_back((int)pl->blocks.size());
    }
    int rc;
    std::vector<int8_t> zero_part(std::max<int64_t>(pl->total_part, 1), 0);
    if ((rc = dev_upload(&pl->d_models, pl->models)) || (rc = dev_upload(&pl->d_blocks, pl->blocks)) ||
        (rc = dev_upload(&pl->d_spans, pl->spans)) || (rc = dev_upload(&pl->d_part, zero_part))) {
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
                                              pl->group_start[g + 1] - pl->group_start[g], pl->group_lds[g],
                                              d_samples, d_edges, pl->d_part, st));
    MISA_HIP(misa::launch_moments_reduce(pl->d_spans, (int)pl->spans.size(), pl->d_part, d_scale, d_grad_h,
                                         d_grad_J, d_sums, st));
    return MISA_OK;
}

}  // extern "C"
