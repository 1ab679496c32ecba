// Test helper: the same push / key-change / pop replay as the oracle's
// orc_heap_replay, on std::push_heap / std::pop_heap over a vector of cell ids
// with a comparator that reads keys live -- the container and algorithms
// boost::heap::priority_queue is built on (fmm/fmdata/fmpriorityqueue.hpp:84,
// comparator FMCompare fmm/fmdata/fmcompare.hpp:29-32).
#include <algorithm>
#include <vector>

extern "C" int heap_replay_std(int ncell, double *keys, int nops, const int *ops,
                               const double *vals, int *popped)
{
    (void)ncell;
    std::vector<int> q;
    auto comp = [keys](int a, int b) { return keys[a] > keys[b]; };
    int np = 0;
    for (int i = 0; i < nops; i++) {
        int kind = ops[2 * i], c = ops[2 * i + 1];
        if (kind == 0) { q.push_back(c); std::push_heap(q.begin(), q.end(), comp); }
        else if (kind == 1) keys[c] = vals[i];
        else if (!q.empty()) {
            popped[np++] = q.front();
            std::pop_heap(q.begin(), q.end(), comp);
            q.pop_back();
        }
    }
    return np;
}
