// Host-only planning of a render call's Halton sample passes into launches ("chunks") and
// lanes. enqueue_passes (c2_device.hip) walks the plan and issues the kernels; the plan
// itself touches no device state, so it is checked on the CPU under AddressSanitizer and
// UndefinedBehaviorSanitizer (tests/sched_check.cpp) for every mix of batch views, sample
// counts, chunk sizes and lane counts.
//
// A pass table lists the views taking part in one sample pass (a single renderer has one
// table for every sample; a batch has table k for sample k: the views with spp > k) and the
// pass's linear index count. A chunk is samples [sample, sample + nsamp) of the linear
// indices [pix_begin, pix_begin + pix_count) of one table. Several samples share a chunk
// only when the chunk holds whole passes whose tables have the same views (the sets shrink
// as k passes each view's spp, so an equal view count means the same set). Chunk c runs on
// lane c % lanes and, with more than one lane, accumulates its film after chunk c - 1's
// accumulation (wait_acc names that chunk's lane), so every pixel's samples are added in
// sample order.
#pragma once
#include <algorithm>
#include <cstdint>
#include <vector>

namespace c2 {

struct SchedPass {
    int32_t off, npv;  // views pv[off, off + npv)
    int64_t npix;      // linear indices of the pass
};

struct SchedChunk {
    int32_t sample, nsamp;          // Halton samples [sample, sample + nsamp)
    int32_t tab;                    // pass table of those samples
    int32_t pix_begin, pix_count;   // linear indices of the chunk (the same for every sample)
    int32_t lane;                   // 0 = the caller's stream
    int32_t wait_acc;               // lane of the previous chunk's film accumulation, -1 none
};

struct SchedInput {
    const SchedPass* passes;
    int32_t npasses;
    bool per_sample_passes;    // table k for sample k (a batch), else table 0 for every sample
    int32_t sample_begin, sample_count;
    int64_t chunk;             // paths the path buffers of one lane hold
    int32_t samples_per_chunk; // at most this many whole passes per chunk (1: one pass, split)
    int32_t lanes;
    // A single renderer's chunks may grow to this many passes (0: samples_per_chunk) so that
    // their count is a multiple of the lanes: with 2 lanes, 20 passes of 4 would leave a
    // fifth chunk running alone, 4 chunks of 5 end together.
    int32_t max_samples_per_chunk;
};

// Samples per chunk of a single renderer's `count` samples: chunks of `spc`, unless their
// count is not a multiple of `lanes`; then the next lower multiple of chunks, the samples
// spread evenly, when no chunk exceeds `cap` samples. Empty: keep chunks of `spc`.
inline std::vector<int64_t> lane_balanced_chunks(int64_t count, int64_t spc, int64_t cap, int32_t lanes) {
    std::vector<int64_t> ns;
    if (count <= 0 || spc < 1 || lanes < 2 || cap <= spc) return ns;
    const int64_t n = (count + spc - 1) / spc;
    const int64_t n2 = n - n % lanes;
    if (n % lanes == 0 || n2 < lanes) return ns;
    if ((count + n2 - 1) / n2 > cap) return ns;
    for (int64_t i = 0; i < n2; i++) ns.push_back(count / n2 + (i < count % n2 ? 1 : 0));
    return ns;
}

// Fills `out` and returns 0, or returns -1 with `out` empty for inputs no plan can serve: a
// negative range, a sample without a pass table, a chunk that would exceed the buffers.
inline int plan_passes_into(const SchedInput& in, std::vector<SchedChunk>& out) {
    out.clear();
    if (!in.passes || in.npasses < 1 || in.lanes < 1 || in.chunk < 1 || in.sample_begin < 0 || in.sample_count < 0 ||
        in.samples_per_chunk < 1)
        return -1;
    const int64_t end = (int64_t)in.sample_begin + in.sample_count;
    if (end > INT32_MAX) return -1;
    std::vector<int64_t> plan;  // a single renderer's chunk sizes, balanced over the lanes
    size_t planned = 0;
    if (!in.per_sample_passes && in.max_samples_per_chunk > in.samples_per_chunk && in.passes[0].npix > 0 &&
        in.passes[0].npix <= INT32_MAX) {
        const int64_t fit = in.chunk / in.passes[0].npix;
        plan = lane_balanced_chunks(in.sample_count, std::min<int64_t>(fit, in.samples_per_chunk),
                                    std::min<int64_t>(fit, in.max_samples_per_chunk), in.lanes);
    }
    int64_t nchunk = 0;
    for (int64_t k = in.sample_begin; k < end;) {
        const int64_t tab = in.per_sample_passes ? k : 0;
        if (tab >= in.npasses) return -1;
        const SchedPass& T = in.passes[tab];
        const int64_t npix = T.npix;
        if (npix < 0 || npix > INT32_MAX) return -1;
        int64_t ns = 1;
        if (in.samples_per_chunk > 1 && npix > 0) {
            int64_t lim = std::min<int64_t>({in.chunk / npix, end - k, (int64_t)in.samples_per_chunk});
            if (in.per_sample_passes)
                for (int64_t j = 1; j < lim; j++) {
                    if (k + j >= in.npasses) return -1;
                    if (in.passes[k + j].npv != T.npv || in.passes[k + j].npix != npix) {
                        lim = j;
                        break;
                    }
                }
            ns = std::max<int64_t>(1, lim);
        }
        if (planned < plan.size()) ns = plan[planned++];
        const int64_t step = ns > 1 ? npix : in.chunk;
        for (int64_t beg = 0; beg < npix; beg += step) {
            const int64_t cnt = std::min<int64_t>(step, npix - beg);
            if (ns * cnt > in.chunk) return -1;
            SchedChunk c;
            c.sample = (int32_t)k;
            c.nsamp = (int32_t)ns;
            c.tab = (int32_t)tab;
            c.pix_begin = (int32_t)beg;
            c.pix_count = (int32_t)cnt;
            c.lane = (int32_t)(nchunk % in.lanes);
            c.wait_acc = (in.lanes > 1 && nchunk > 0) ? (int32_t)((nchunk - 1) % in.lanes) : -1;
            out.push_back(c);
            nchunk++;
        }
        k += ns;
    }
    return 0;
}

inline int plan_passes(const SchedInput& in, std::vector<SchedChunk>& out) {
    const int rc = plan_passes_into(in, out);
    if (rc != 0) out.clear();
    return rc;
}

}  // namespace c2
