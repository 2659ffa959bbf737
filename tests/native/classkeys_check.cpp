// Host check of GridInterpolationBackward's sort keys (det_scatter.h
// ClassKeys): over every cell of small even grids, key -> cell inverts
// cell -> key, the keys are a permutation of [0, R^3) (so R^3 is free for
// invalid samples), a key's range [c h^3, (c + 1) h^3) names the cell's parity
// class c, and within a class the keys rise with the cell index (the order
// the run sums need).  Large grids (up to 1290^3 cells) are checked on random
// cells and the grid's corners.  Prints "ok <cells>" or the first failure.
#include "det_scatter.h"

#include <cstdio>
#include <random>
#include <vector>

using namespace qsdvr;

static int check_cell(const ClassKeys &ck, int64_t x, int64_t y, int64_t z, uint32_t &k) {
    const int64_t R = ck.R, cell = x + R * (y + R * z);
    k = ck.key(x, y, z);
    const uint32_t cls = (uint32_t)((x & 1) | ((y & 1) << 1) | ((z & 1) << 2));
    if (ck.cell(k) != cell || (uint64_t)k >= (uint64_t)R * R * R || k / ck.h3 != cls) {
        printf("R=%lld cell (%lld %lld %lld): key %u decodes to %lld, class %u\n", (long long)R,
               (long long)x, (long long)y, (long long)z, k, (long long)ck.cell(k), k / ck.h3);
        return 1;
    }
    return 0;
}

int main() {
    long long cells = 0;
    for (int R : {2, 4, 6, 10, 16, 34, 128}) {
        const ClassKeys ck = class_keys(R);
        if ((uint64_t)8 * ck.h3 != (uint64_t)R * R * R) {
            printf("R=%d: 8 h^3 != R^3\n", R);
            return 1;
        }
        std::vector<char> seen((size_t)R * R * R, 0);
        std::vector<int64_t> last(8, -1);
        for (int64_t z = 0; z < R; ++z)
            for (int64_t y = 0; y < R; ++y)
                for (int64_t x = 0; x < R; ++x) {
                    uint32_t k;
                    if (check_cell(ck, x, y, z, k)) return 1;
                    if (seen[k]) {
                        printf("R=%d: key %u taken twice\n", R, k);
                        return 1;
                    }
                    seen[k] = 1;
                    const uint32_t cls = k / ck.h3;
                    if ((int64_t)k <= last[cls]) {
                        printf("R=%d: class %u keys not rising with the cell index\n", R, cls);
                        return 1;
                    }
                    last[cls] = k;
                    ++cells;
                }
    }
    std::mt19937_64 rng(7);
    for (int R : {130, 512, 1024, 1290}) {
        const ClassKeys ck = class_keys(R);
        std::uniform_int_distribution<int64_t> u(0, R - 1);
        for (int i = 0; i < 200000; ++i) {
            uint32_t k;
            if (check_cell(ck, u(rng), u(rng), u(rng), k)) return 1;
            ++cells;
        }
        for (int c = 0; c < 8; ++c) {
            uint32_t k;
            if (check_cell(ck, (c & 1) ? R - 1 : 0, (c & 2) ? R - 1 : 0, (c & 4) ? R - 1 : 0, k))
                return 1;
            ++cells;
        }
    }
    printf("ok %lld\n", cells);
    return 0;
}
