// qiw_chunks.h -- where the host compiler cuts a vector program into topology chunks.
// Plain C++ (no HIP): the CPU tests compile it on its own.
#pragma once
#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <vector>

// The first topology of each chunk, for topologies of costs tcost (the compiler's words x
// columns) taken in order.  A chunk closes at a topology once it has reached chunk_cost
// (chunk_cost <= 0: one chunk).  With multiple > 1 (the matrix-core walk runs the chunks of
// one sample block `multiple` to a workgroup, which lasts as long as its longest chunk) that
// count is rounded up to a multiple of `multiple`, as far as the topologies go, and the
// chunks are cut under the smallest cap on a chunk's cost that packs the topologies, in
// order, into that many chunks: no chunk costs more than the longest of the first rule.
inline std::vector<size_t> qiw_chunk_starts(const std::vector<int64_t> &tcost, int64_t chunk_cost, int multiple) {
    const size_t nt = tcost.size();
    // cap > 0: close before a chunk would exceed cap; cap == 0: close once it reaches chunk_cost
    auto cut = [&](int64_t cap, std::vector<size_t> *starts) {
        int64_t n = 1, cur = 0;
        if (starts) starts->assign(1, 0);
        for (size_t t = 0, t0 = 0; t < nt; ++t) {
            if (chunk_cost > 0 && t > t0 && (cap > 0 ? cur + tcost[t] > cap : cur >= chunk_cost)) {
                ++n;
                cur = 0;
                t0 = t;
                if (starts) starts->push_back(t);
            }
            cur += tcost[t];
        }
        return n;
    };
    int64_t cap = 0;
    if (chunk_cost > 0 && multiple > 1 && nt > 1) {
        int64_t target = cut(0, nullptr);
        target = (target + multiple - 1) / multiple * multiple;
        target = std::min<int64_t>(target, (int64_t)nt);
        int64_t total = 0;
        for (int64_t c : tcost) total += c;
        int64_t lo = *std::max_element(tcost.begin(), tcost.end()), hi = std::max<int64_t>(total, 1);
        while (lo < hi) {
            const int64_t mid = lo + (hi - lo) / 2;
            if (cut(mid, nullptr) <= target)
                hi = mid;
            else
                lo = mid + 1;
        }
        cap = std::max<int64_t>(lo, 1);
    }
    std::vector<size_t> starts;
    cut(cap, &starts);
    return starts;
}
