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
    std::vector<long long> len(nnz + 1, 0);
    std::vector<std::thread> th;
    std::vector<int> k0(nt + 1);
    for (int w = 0; w <= nt; ++w) k0[w] = (int)((long long)n * w / nt);
    for (int w = 0; w < nt; ++w)
        th.emplace_back(planRows, rs, lend, ubeg, col, k0[w], k0[w + 1], len.data(), std::ref(pr[w]));
    for (auto& x : th) x.join();
    start.assign(nnz + 1, 0);
    for (long long a = 0; a < nnz; ++a) start[a + 1] = start[a] + len[a];
    prog.clear();
    prog.reserve((size_t)start[nnz]);
    for (int w = 0; w < nt; ++w) {
        prog.insert(prog.end(), pr[w].begin(), pr[w].end());
        std::vector<int>().swap(pr[w]);
    }
}

}  // namespace astg
