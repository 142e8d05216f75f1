// CPU check of the render planner (crystal2_amd/csrc/c2_sched.h), built with
// AddressSanitizer and UndefinedBehaviorSanitizer by tests/test_sched.py. Random single
// renders and batches (views of uneven pixel counts and spp, pass tables built the way the
// renderer builds them), chunk sizes, samples per chunk and 1..4 lanes; for every plan:
//   * every (sample, linear index) of the range is covered exactly once;
//   * a chunk holds at most `chunk` paths, and its samples' pass tables list the same views;
//   * chunk c runs on lane c % lanes and waits for chunk c - 1's accumulation;
//   * each pixel of each view receives its samples in increasing order (the film adds
//     follow the plan order);
// and inputs no plan can serve are refused.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "c2_sched.h"

using namespace c2;

namespace {

int failures = 0;
#define CHECK(cond, ...)                                  \
    do {                                                  \
        if (!(cond)) {                                    \
            if (failures++ < 20) {                        \
                std::fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
                std::fprintf(stderr, __VA_ARGS__);        \
                std::fprintf(stderr, "\n");               \
            }                                             \
        }                                                 \
    } while (0)

struct Config {
    std::vector<int64_t> view_pix;  // linear indices (pixels) of each view
    std::vector<int32_t> view_spp;
    bool batch;
};

// Pass tables as pt_build makes them: a batch has table k with the views of spp > k (in view
// order, each view's pixels after the previous one's); a single renderer has one table.
void build_tables(const Config& c, std::vector<SchedPass>& tabs, std::vector<std::vector<int>>& views) {
    tabs.clear();
    views.clear();
    int32_t npasses = 1;
    if (c.batch)
        for (int32_t s : c.view_spp) npasses = std::max(npasses, s);
    int32_t off = 0;
    for (int32_t k = 0; k < npasses; k++) {
        SchedPass t{off, 0, 0};
        std::vector<int> vs;
        for (size_t v = 0; v < c.view_pix.size(); v++) {
            if (c.batch && c.view_spp[v] <= k) continue;
            vs.push_back((int)v);
            t.npix += c.view_pix[v];
            t.npv++;
        }
        off += t.npv;
        tabs.push_back(t);
        views.push_back(vs);
    }
}

void check_plan(const Config& c, const SchedInput& in, const std::vector<SchedPass>& tabs,
                const std::vector<std::vector<int>>& views, const std::vector<SchedChunk>& plan) {
    const int64_t end = (int64_t)in.sample_begin + in.sample_count;
    // coverage per sample: one count per linear index
    std::vector<std::vector<uint8_t>> cover;
    for (int64_t k = in.sample_begin; k < end; k++)
        cover.emplace_back((size_t)tabs[c.batch ? k : 0].npix, 0);
    // last sample added per (view, pixel)
    std::vector<std::vector<int32_t>> last(c.view_pix.size());
    for (size_t v = 0; v < c.view_pix.size(); v++) last[v].assign((size_t)c.view_pix[v], -1);
    for (size_t i = 0; i < plan.size(); i++) {
        const SchedChunk& ch = plan[i];
        CHECK(ch.lane == (int32_t)(i % in.lanes), "chunk %zu lane %d", i, ch.lane);
        CHECK(ch.wait_acc == (in.lanes > 1 && i > 0 ? (int32_t)((i - 1) % in.lanes) : -1), "chunk %zu wait %d", i,
              ch.wait_acc);
        CHECK((int64_t)ch.nsamp * ch.pix_count <= in.chunk, "chunk %zu holds %lld paths > %lld", i,
              (long long)ch.nsamp * ch.pix_count, (long long)in.chunk);
        CHECK(ch.nsamp >= 1 && ch.pix_count >= 1 && ch.pix_begin >= 0, "chunk %zu empty", i);
        CHECK(ch.tab >= 0 && ch.tab < (int32_t)tabs.size(), "chunk %zu table %d", i, ch.tab);
        if (ch.tab < 0 || ch.tab >= (int32_t)tabs.size()) continue;
        CHECK(ch.tab == (c.batch ? ch.sample : 0), "chunk %zu table %d for sample %d", i, ch.tab, ch.sample);
        const SchedPass& T = tabs[ch.tab];
        CHECK((int64_t)ch.pix_begin + ch.pix_count <= T.npix, "chunk %zu beyond its pass", i);
        for (int32_t s = ch.sample; s < ch.sample + ch.nsamp; s++) {
            CHECK(s >= in.sample_begin && s < end, "chunk %zu sample %d outside the range", i, s);
            if (s < in.sample_begin || s >= end) continue;
            const int32_t ts = c.batch ? s : 0;
            CHECK(views[ts] == views[ch.tab], "chunk %zu: sample %d has other views than its table", i, s);
            auto& cv = cover[s - in.sample_begin];
            for (int64_t li = ch.pix_begin; li < (int64_t)ch.pix_begin + ch.pix_count && li < (int64_t)cv.size(); li++) {
                cv[li]++;
                // the view and pixel of linear index li in table ts
                int64_t rem = li;
                for (int v : views[ts]) {
                    if (rem < c.view_pix[v]) {
                        int32_t& l = last[v][rem];
                        CHECK(s > l, "view %d pixel %lld: sample %d after %d", v, (long long)rem, s, l);
                        l = s;
                        break;
                    }
                    rem -= c.view_pix[v];
                }
            }
        }
    }
    for (size_t k = 0; k < cover.size(); k++)
        for (size_t li = 0; li < cover[k].size(); li++)
            CHECK(cover[k][li] == 1, "sample %lld index %zu covered %d times", (long long)(in.sample_begin + (int64_t)k),
                  li, cover[k][li]);
}

}  // namespace

int main() {
    std::mt19937 rng(2024);
    auto U = [&](int a, int b) { return std::uniform_int_distribution<int>(a, b)(rng); };
    int plans = 0, chunks = 0;
    for (int trial = 0; trial < 240; trial++) {
        Config c;
        c.batch = trial % 3 != 0;
        const int nv = c.batch ? U(1, 12) : 1;
        for (int v = 0; v < nv; v++) {
            c.view_pix.push_back(U(1, 3000));
            const int pick = U(0, 3);
            c.view_spp.push_back(pick == 0 ? 4 : pick == 1 ? 16 : pick == 2 ? U(1, 20) : 1);
        }
        std::vector<SchedPass> tabs;
        std::vector<std::vector<int>> views;
        build_tables(c, tabs, views);
        int64_t maxPass = 0;
        for (auto& t : tabs) maxPass = std::max(maxPass, t.npix);
        const int32_t spp = c.batch ? (int32_t)tabs.size() : c.view_spp[0] + U(0, 6);
        for (int lanes = 1; lanes <= 4; lanes++)
            for (int spc : {1, 3, 4, 16}) {
                const int64_t chunks_opts[] = {(int64_t)U(16, 97), (int64_t)U(200, 900), maxPass, maxPass * U(1, 5) + U(0, 50),
                                               std::max<int64_t>(1, maxPass / U(2, 7))};
                for (int64_t chunk : chunks_opts) {
                    SchedInput in{};
                    in.passes = tabs.data();
                    in.npasses = (int32_t)tabs.size();
                    in.per_sample_passes = c.batch;
                    in.sample_begin = c.batch ? (U(0, 3) == 0 ? U(0, spp - 1) : 0) : U(0, 5);
                    in.sample_count = c.batch ? spp - in.sample_begin : U(0, spp);
                    in.chunk = chunk;
                    in.samples_per_chunk = spc;
                    in.lanes = lanes;
                    // single renderers: chunks may grow by half to end on a multiple of the lanes
                    in.max_samples_per_chunk = (!c.batch && U(0, 1)) ? spc + spc / 2 : 0;
                    std::vector<SchedChunk> plan;
                    const int rc = plan_passes(in, plan);
                    CHECK(rc == 0, "trial %d: plan refused", trial);
                    if (rc != 0) continue;
                    check_plan(c, in, tabs, views, plan);
                    plans++;
                    chunks += (int)plan.size();
                }
            }
        // inputs no plan can serve
        if (c.batch) {
            SchedInput in{};
            in.passes = tabs.data();
            in.npasses = (int32_t)tabs.size();
            in.per_sample_passes = true;
            in.sample_begin = 0;
            in.sample_count = (int32_t)tabs.size() + 1;  // one sample beyond the tables
            in.chunk = maxPass;
            in.samples_per_chunk = 1;
            in.lanes = 2;
            std::vector<SchedChunk> plan;
            CHECK(plan_passes(in, plan) == -1 && plan.empty(), "trial %d: a sample beyond the tables was planned",
                  trial);
        }
        SchedInput bad{};
        bad.passes = tabs.data();
        bad.npasses = (int32_t)tabs.size();
        bad.sample_begin = -1;
        bad.sample_count = 1;
        bad.chunk = 10;
        bad.samples_per_chunk = 1;
        bad.lanes = 1;
        std::vector<SchedChunk> plan;
        CHECK(plan_passes(bad, plan) == -1, "a negative range was planned");
        bad.sample_begin = 0;
        bad.lanes = 0;
        CHECK(plan_passes(bad, plan) == -1, "zero lanes were planned");
    }
    {  // a single renderer's chunk sizes with chunks allowed to grow from 4 to 6 samples
        std::vector<SchedPass> tab{{0, 1, 1000}};
        struct Case {
            int count, lanes;
            std::vector<int> want;
        };
        const Case cases[] = {{20, 2, {5, 5, 5, 5}},     {16, 2, {4, 4, 4, 4}}, {12, 2, {6, 6}},
                              {4, 2, {4}},               {9, 2, {5, 4}},        {13, 2, {4, 4, 4, 1}},
                              {28, 2, {5, 5, 5, 5, 4, 4}}, {20, 3, {4, 4, 4, 4, 4}}, {20, 1, {4, 4, 4, 4, 4}},
                              {18, 3, {6, 6, 6}}};
        for (const Case& cs : cases) {
            SchedInput in{};
            in.passes = tab.data();
            in.npasses = 1;
            in.sample_begin = 5;
            in.sample_count = cs.count;
            in.chunk = 1000 * 6;
            in.samples_per_chunk = 4;
            in.max_samples_per_chunk = 6;
            in.lanes = cs.lanes;
            std::vector<SchedChunk> plan;
            CHECK(plan_passes(in, plan) == 0, "plan of %d samples refused", cs.count);
            std::vector<int> got;
            int next = in.sample_begin;
            for (const SchedChunk& ch : plan) {
                got.push_back(ch.nsamp);
                CHECK(ch.sample == next, "chunk starts at sample %d, expected %d", ch.sample, next);
                next += ch.nsamp;
            }
            CHECK(got == cs.want, "chunk sizes of %d samples on %d lanes", cs.count, cs.lanes);
        }
    }
    std::printf("sched_check: %d plans, %d chunks, %d failures\n", plans, chunks, failures);
    return failures ? 1 : 0;
}
