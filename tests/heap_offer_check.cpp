// lightCache_t::gatherSamples' offer of one candidate into its found heap
// (lightcache.cc:147-157) with the real std::push_heap / std::pop_heap of
// the C++ library: reads "limit n w_0 .. w_{n-1}" per line from stdin and
// prints the final heap array as "id:weight" pairs, one line per input.
#include <algorithm>
#include <cstdio>
#include <vector>

struct Found {
    int S;
    float dis, weight;
};
struct CompareFound {
    bool operator()(const Found& a, const Found& b) const { return a.weight > b.weight; }
};

int main()
{
    unsigned limit;
    int n;
    while (std::scanf("%u %d", &limit, &n) == 2) {
        std::vector<Found> found;
        for (int i = 0; i < n; i++) {
            float w;
            if (std::scanf("%f", &w) != 1) return 1;
            Found t{i, 0.f, w};
            if (found.size() >= limit && t.weight < found.front().weight) continue;
            found.push_back(t);
            std::push_heap(found.begin(), found.end(), CompareFound());
            if (found.size() > limit) {
                std::pop_heap(found.begin(), found.end(), CompareFound());
                found.pop_back();
            }
        }
        for (const Found& f : found) std::printf("%d:%.9g ", f.S, f.weight);
        std::printf("\n");
    }
    return 0;
}
