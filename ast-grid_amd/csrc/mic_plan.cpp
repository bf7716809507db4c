// Symbolic plan of the MIC(0) factorisation (see mic_plan.h).
#include "mic_plan.h"

#include <algorithm>
#include <thread>

namespace astg {

namespace {

// Rows [k0, k1): segment lengths into start[] (by position), data into prog.
void planRows(const int* rs, const int* lend, const int* ubeg, const int* col, int k0, int k1, long long* len,
              std::vector<int>& prog) {
    for (int k = k0; k < k1; ++k) {
        const int r0 = rs[k], le = lend[k], ub = ubeg[k], r1 = rs[k + 1];
        for (int a = r0; a < r1; ++a) len[a] = 0;
        for (int a = r0; a < le; ++a) {
            const int m = col[a];
            int b = r0, cnt = 0;
            for (int p = ubeg[m]; p < rs[m + 1]; ++p) {
                const int c = col[p];
                if (c == k) continue;
                while (b < r1 && col[b] < c) ++b;
                if (b < r1 && col[b] == c) continue;
                prog.push_back(p);
                ++cnt;
            }
            len[a] = cnt;
        }
        for (int p = ub; p < r1; ++p) {
            const int i = col[p];
            int a = r0, c = rs[i], cnt = 0;
            const int ce = lend[i];
            while (a < le && c < ce) {
                if (col[a] < col[c]) {
                    ++a;
                } else if (col[c] < col[a]) {
                    ++c;
                } else {
                    prog.push_back(a - r0);
                    prog.push_back(c);
                    cnt += 2;
                    ++a;
                    ++c;
                }
            }
            len[p] = cnt;
        }
    }
}

}  // namespace

void planMic(int n, const int* rs, const int* lend, const int* ubeg, const int* col, std::vector<long long>& start,
             std::vector<int>& prog) {
    const long long nnz = rs[n];
    const int nt = std::max(1, std::min(16, (int)std::thread::hardware_concurrency()));
    std::vector<std::vector<int>> pr(nt);
    start.resize(nnz + 1);
    long long* len = start.data() + 1;  // segment lengths first, prefix sums in place afterwards
    std::vector<int> k0(nt + 1);
    for (int w = 0; w <= nt; ++w) k0[w] = (int)((long long)n * w / nt);
    {
        std::vector<std::thread> th;
        for (int w = 0; w < nt; ++w)
            th.emplace_back(planRows, rs, lend, ubeg, col, k0[w], k0[w + 1], len, std::ref(pr[w]));
        for (auto& x : th) x.join();
    }
    // thread w's segments are contiguous (rows k0[w]..k0[w+1]), so its prefix
    // sums start at the total of the threads before it
    std::vector<long long> base(nt + 1, 0);
    for (int w = 0; w < nt; ++w) base[w + 1] = base[w] + (long long)pr[w].size();
    start[0] = 0;
    prog.resize((size_t)base[nt]);
    {
        std::vector<std::thread> th;
        for (int w = 0; w < nt; ++w)
            th.emplace_back([&, w] {
                long long acc = base[w];
                for (long long a = rs[k0[w]]; a < rs[k0[w + 1]]; ++a) {
                    acc += len[a];
                    len[a] = acc;  // start[a + 1]
                }
                std::copy(pr[w].begin(), pr[w].end(), prog.begin() + base[w]);
                std::vector<int>().swap(pr[w]);
            });
        for (auto& x : th) x.join();
    }
}

}  // namespace astg
