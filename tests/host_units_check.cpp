// Host check of the runtime's plan arithmetic, built with g++ by
// tests/test_host_units.py: the units of a cancellable call (plan_unit)
// partition the plan's bands in order, each unit is exactly the union of
// its bands (so finishing a prefix of units finishes a prefix of bands,
// internal/exec/doc.go:165-183), and a target above the raster's cells
// gives one unit.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "plan.hpp"

using namespace strata;

static int failures = 0;
#define CHECK(c, ...)                                                                                        \
    do {                                                                                                     \
        if (!(c)) {                                                                                          \
            if (failures++ < 20) {                                                                           \
                fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__);                                         \
                fprintf(stderr, __VA_ARGS__);                                                                \
                fprintf(stderr, "\n");                                                                       \
            }                                                                                                \
        }                                                                                                    \
    } while (0)

static int64_t area(const Rect &r) { return (r.x1 - r.x0) * (r.y1 - r.y0); }

static void check_units(int64_t W, int64_t H, int64_t tw, int64_t th, int64_t target)
{
    const Plan p = new_plan(W, H, tw, th);
    int64_t next = 0, units = 0;
    while (next < p.bands) {
        const Unit u = plan_unit(p, next, target);
        CHECK(u.b0 == next, "W%lld H%lld t%lldx%lld target %lld: unit starts at %lld, expected %lld", (long long)W,
              (long long)H, (long long)tw, (long long)th, (long long)target, (long long)u.b0, (long long)next);
        CHECK(u.b1 > u.b0 && u.b1 <= p.bands, "unit [%lld, %lld) of %lld bands", (long long)u.b0, (long long)u.b1,
              (long long)p.bands);
        CHECK(u.rc.x0 >= 0 && u.rc.y0 >= 0 && u.rc.x1 <= W && u.rc.y1 <= H && u.rc.x0 < u.rc.x1 && u.rc.y0 < u.rc.y1,
              "unit rect outside the raster");
        int64_t sum = 0;
        for (int64_t i = u.b0; i < u.b1 && i < p.bands; ++i) {
            const Rect b = plan_band(p, i);
            sum += area(b);
            CHECK(b.x0 >= u.rc.x0 && b.x1 <= u.rc.x1 && b.y0 >= u.rc.y0 && b.y1 <= u.rc.y1,
                  "W%lld H%lld t%lldx%lld target %lld: band %lld outside its unit", (long long)W, (long long)H,
                  (long long)tw, (long long)th, (long long)target, (long long)i);
        }
        // Bands never overlap, so equal areas mean the bands fill the rect.
        CHECK(sum == area(u.rc), "W%lld H%lld t%lldx%lld target %lld: unit %lld covers %lld cells, its bands %lld",
              (long long)W, (long long)H, (long long)tw, (long long)th, (long long)target, (long long)units,
              (long long)area(u.rc), (long long)sum);
        next = u.b1;
        ++units;
    }
    CHECK(next == p.bands, "units end at %lld of %lld bands", (long long)next, (long long)p.bands);
    if (target > W * H) CHECK(units == 1, "a target above the raster gave %lld units", (long long)units);
}

int main()
{
    const int64_t sizes[][2] = {{1, 1}, {1, 5}, {5, 1}, {7, 3}, {64, 64}, {301, 97}, {1000, 1},
                                {1, 1000}, {20000, 7}, {4096, 4096}, {65537, 3}, {513, 1029}};
    const int64_t tiles[][2] = {{0, 0}, {1, 0}, {7, 5}, {64, 1}, {256, 256}, {1000, 3}, {0, 16}, {3, 1000}};
    const int64_t targets[] = {1, 100, 5000, 1 << 16, 1 << 20, 1 << 22, (int64_t)1 << 40};
    for (auto &s : sizes)
        for (auto &t : tiles)
            for (int64_t target : targets) check_units(s[0], s[1], t[0], t[1], target);
    if (failures) {
        fprintf(stderr, "%d failures\n", failures);
        return 1;
    }
    printf("host units: ok\n");
    return 0;
}
