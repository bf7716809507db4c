// Level merging for the sync-free triangular sweeps: symbolic plan.
#include "sweep_merge.h"

#include <algorithm>
#include <thread>

namespace astg {

namespace {

struct Part {
    std::vector<int> start, nr, idx, ustart;
    std::vector<uint8_t> upd;
    long long rTerms = 0, yTerms = 0, kept = 0;
};

int find(const std::vector<int>& v, int k) {
    for (size_t a = 0; a < v.size(); ++a)
        if (v[a] == k) return (int)a;
    return -1;
}

void planRange(int dir, const int* rs, const int* lend, const int* ubeg, const int* col, const int* lev, int q,
               int maxY, const int* order, const int* pos, int t0, int t1, Part& out) {
    auto d0 = [&](int x) { return dir == 0 ? rs[x] : ubeg[x]; };
    auto d1 = [&](int x) { return dir == 0 ? lend[x] : rs[x + 1]; };
    std::vector<int> keys, state, sub, ys, slot;
    out.start.reserve(t1 - t0);
    for (int t = t0; t < t1; ++t) {
        out.start.push_back((int)out.idx.size());
        out.ustart.push_back((int)out.upd.size());
        const int i = order[t];
        if (i < 0) {
            out.nr.push_back(0);
            continue;
        }
        const int g = lev[i] / q;
        keys.assign(col + d0(i), col + d1(i));
        state.assign(keys.size(), 0);  // 0 open, 1 substituted, 2 kept
        sub.clear();
        int live = (int)keys.size();
        for (;;) {
            int best = -1;
            for (size_t a = 0; a < keys.size(); ++a) {
                if (state[a] != 0 || lev[keys[a]] / q != g) continue;
                if (best < 0 || (dir == 0 ? lev[keys[a]] > lev[keys[best]] : lev[keys[a]] < lev[keys[best]]))
                    best = (int)a;
            }
            if (best < 0) break;
            const int j = keys[best];
            int add = 0;
            for (int p = d0(j); p < d1(j); ++p)
                if (find(keys, col[p]) < 0) ++add;
            if (live - 1 + add > maxY || (int)sub.size() >= kMaxRTerms) {
                state[best] = 2;
                ++out.kept;
                continue;
            }
            state[best] = 1;
            sub.push_back(j);
            live += add - 1;
            for (int p = d0(j); p < d1(j); ++p)
                if (find(keys, col[p]) < 0) {
                    keys.push_back(col[p]);
                    state.push_back(0);
                }
        }
        ys.clear();
        for (size_t a = 0; a < keys.size(); ++a)
            if (state[a] != 1) ys.push_back(keys[a]);
        std::sort(ys.begin(), ys.end());
        // slot of every key: substituted keys first (in order), then the y-terms
        slot.assign(keys.size(), -1);
        for (size_t a = 0; a < keys.size(); ++a) {
            const int k = keys[a];
            if (state[a] == 1) {
                slot[a] = (int)(std::find(sub.begin(), sub.end(), k) - sub.begin());
            } else {
                slot[a] = (int)sub.size() + (int)(std::lower_bound(ys.begin(), ys.end(), k) - ys.begin());
            }
        }
        auto slotOf = [&](int k) { return (uint8_t)slot[find(keys, k)]; };
        for (int p = d0(i); p < d1(i); ++p) out.upd.push_back(slotOf(col[p]));
        for (int j : sub)
            for (int p = d0(j); p < d1(j); ++p) out.upd.push_back(slotOf(col[p]));
        for (int j : sub) out.idx.push_back(dir == 0 ? j : pos[j]);
        for (int k : ys) out.idx.push_back(pos[k]);
        out.nr.push_back((int)sub.size());
        out.rTerms += (long long)sub.size();
        out.yTerms += (long long)ys.size();
    }
}

}  // namespace

void planStencil(int dir, int n, const int* rs, const int* lend, const int* ubeg, const int* col, const int* lev,
                 int q, int maxY, const int* order, int P, const int* pos, SweepStencil& st) {
    (void)n;
    q = std::max(1, q);
    const int nt = std::max(1, std::min(16, (int)std::thread::hardware_concurrency()));
    std::vector<Part> parts(nt);
    std::vector<std::thread> th;
    const int per = (P + nt - 1) / nt;
    for (int w = 0; w < nt; ++w) {
        const int t0 = std::min(P, w * per), t1 = std::min(P, t0 + per);
        th.emplace_back(planRange, dir, rs, lend, ubeg, col, lev, q, maxY, order, pos, t0, t1, std::ref(parts[w]));
    }
    for (auto& x : th) x.join();
    // concatenate the parts in parallel at their final offsets
    std::vector<size_t> pS(nt + 1, 0), pE(nt + 1, 0), pU(nt + 1, 0);
    st.rTerms = st.yTerms = st.kept = 0;
    for (int w = 0; w < nt; ++w) {
        pS[w + 1] = pS[w] + parts[w].start.size();
        pE[w + 1] = pE[w] + parts[w].idx.size();
        pU[w + 1] = pU[w] + parts[w].upd.size();
        st.rTerms += parts[w].rTerms;
        st.yTerms += parts[w].yTerms;
        st.kept += parts[w].kept;
    }
    st.start.resize(pS[nt] + 1);
    st.ustart.resize(pS[nt] + 1);
    st.nr.resize(pS[nt]);
    st.idx.resize(pE[nt]);
    st.upd.resize(pU[nt]);
    th.clear();
    for (int w = 0; w < nt; ++w)
        th.emplace_back([&, w] {
            Part& pt = parts[w];
            for (size_t a = 0; a < pt.start.size(); ++a) {
                st.start[pS[w] + a] = (int)pE[w] + pt.start[a];
                st.ustart[pS[w] + a] = (int)pU[w] + pt.ustart[a];
            }
            std::copy(pt.nr.begin(), pt.nr.end(), st.nr.begin() + pS[w]);
            std::copy(pt.idx.begin(), pt.idx.end(), st.idx.begin() + pE[w]);
            std::copy(pt.upd.begin(), pt.upd.end(), st.upd.begin() + pU[w]);
            pt = Part();
        });
    for (auto& x : th) x.join();
    st.start[pS[nt]] = (int)st.idx.size();
    st.ustart[pS[nt]] = (int)st.upd.size();
}

}  // namespace astg
