// Host check of the closed-form ray cut (raygen.h cut_closed_form) against
// the reference's walk over every candidate step (cut_walk): counts and
// first inside steps must be bit-identical.  Rays: random cameras and
// pixels, axis-aligned rays, rays through cube corners and edges, and rays
// grazing the faces.  Prints "ok <rays>" or the first mismatch.
#include "raygen.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>

using namespace qsdvr;

static bool same(f3 a, f3 b) { return memcmp(&a, &b, sizeof(f3)) == 0; }

static int check(f3 n, f3 dir, float interval, uint32_t cand, long long &rays) {
    f3 o1, o2;
    const uint32_t c1 = cut_walk(n, dir, interval, cand, o1);
    const uint32_t c2 = cut_closed_form(n, dir, interval, cand, o2);
    ++rays;
    if (c1 != c2 || !same(o1, o2)) {
        printf("mismatch: n=(%.9g %.9g %.9g) dir=(%.9g %.9g %.9g) interval=%.9g cand=%u walk=%u "
               "closed=%u\n", n.x, n.y, n.z, dir.x, dir.y, dir.z, interval, cand, c1, c2);
        return 1;
    }
    return 0;
}

int main(int argc, char **argv) {
    const long long n_random = argc > 1 ? atoll(argv[1]) : 2000000;
    std::mt19937_64 g(12345);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    long long rays = 0;
    const float intervals[] = {0.027063f, 0.018042f, 0.013532f, 0.009021f, 0.25f, 0.1f};
    // random rays from points around the cube towards random targets
    for (long long i = 0; i < n_random; ++i) {
        const float interval = intervals[i % 6];
        double o[3], t[3], d[3];
        for (int a = 0; a < 3; ++a) { o[a] = 3.5 * U(g); t[a] = 1.2 * U(g); }
        double len = 0;
        for (int a = 0; a < 3; ++a) { d[a] = t[a] - o[a]; len += d[a] * d[a]; }
        len = std::sqrt(len);
        if (len < 1e-3) continue;
        f3 n{(float)o[0], (float)o[1], (float)o[2]};
        f3 dir{(float)(d[0] / len), (float)(d[1] / len), (float)(d[2] / len)};
        // sometimes snap components to 0 or tiny values (axis-aligned, grazing)
        const int mode = (int)(i % 7);
        if (mode == 1) dir.y = 0.0f;
        if (mode == 2) { dir.y = 0.0f; dir.z = 0.0f; dir.x = dir.x < 0 ? -1.0f : 1.0f; }
        if (mode == 3) dir.z = (float)(1e-6 * U(g));
        if (mode == 4) n.y = (i & 8) ? 0.9999f : -0.99989998f;  // on / just inside a face plane
        if (mode == 5) { dir.x = (float)(3e-5 * U(g)); n.x = 0.99985f; }
        const uint32_t cand = (uint32_t)(8.0 / interval);
        if (check(n, dir, interval, cand, rays)) return 1;
    }
    // rays through the cube's corners and edge midpoints from far away
    for (int c = 0; c < 27; ++c) {
        const double p[3] = {(double)(c % 3 - 1) * 0.9999, (double)((c / 3) % 3 - 1) * 0.9999,
                             (double)(c / 9 - 1) * 0.9999};
        for (int k = 0; k < 2000; ++k) {
            double o[3], d[3], len = 0;
            for (int a = 0; a < 3; ++a) o[a] = 3.0 * U(g);
            for (int a = 0; a < 3; ++a) { d[a] = p[a] - o[a]; len += d[a] * d[a]; }
            len = std::sqrt(len);
            f3 n{(float)o[0], (float)o[1], (float)o[2]};
            f3 dir{(float)(d[0] / len), (float)(d[1] / len), (float)(d[2] / len)};
            for (float interval : intervals)
                if (check(n, dir, interval, (uint32_t)(8.0 / interval), rays)) return 1;
        }
    }
    printf("ok %lld\n", rays);
    return 0;
}
