// CPU check of the autotuner's rule (parallel-ra_amd/csrc/pra_tune.hpp):
// variants take turns until each has its samples, then the cheapest mean is
// kept for good; bands, knobs and devices are separate.  Prints "ok" and exits
// 0 when every check holds; otherwise names the first failed check.
#include "../parallel-ra_amd/csrc/pra_tune.hpp"

#include <cstdio>

namespace {
int failures = 0;
void expect(bool ok, const char* what)
{
    if (!ok) {
        std::printf("FAIL: %s\n", what);
        failures++;
    }
}
} // namespace

int main()
{
    pra::Tuner t(3);
    // three variants with true costs 5, 2, 9 per unit (plus noise)
    const double cost[3] = {5.0, 2.0, 9.0};
    int turns[3] = {0, 0, 0};
    for (int i = 0; i < 9; i++) {
        const int v = t.pick(0, 0, 20, 3);
        expect(t.settled(0, 0, 20) == -1, "not settled while variants lack samples");
        turns[v]++;
        t.record(0, 0, 20, v, cost[v] + 0.1 * (i % 3), 3);
    }
    expect(turns[0] == 3 && turns[1] == 3 && turns[2] == 3, "variants take turns, three samples each");
    expect(t.settled(0, 0, 20) == 1, "the cheapest variant is kept");
    expect(t.pick(0, 0, 20, 3) == 1, "pick returns the settled variant");
    t.record(0, 0, 20, 2, 0.0, 3);
    expect(t.pick(0, 0, 20, 3) == 1, "a settled choice does not change");

    // another band of the same knob, another knob, another device: untouched
    expect(t.settled(0, 0, 21) == -1 && t.pick(0, 0, 21, 3) == 0, "bands are separate");
    expect(t.settled(0, 1, 20) == -1, "knobs are separate");
    expect(t.settled(1, 0, 20) == -1, "devices are separate");

    // out-of-range records are ignored
    t.record(0, 1, 5, 7, 1.0, 2);
    t.record(0, 1, 5, 0, -1.0, 2);
    expect(t.pick(0, 1, 5, 2) == 0, "invalid records leave the counts alone");

    // two variants where the second is cheaper: settles after 2 x 3 samples
    for (int i = 0; i < 6; i++) {
        const int v = t.pick(0, 1, 5, 2);
        t.record(0, 1, 5, v, v == 0 ? 3.0 : 1.0, 2);
    }
    expect(t.settled(0, 1, 5) == 1, "two-variant knob settles on the cheaper");

    const char* names[2] = {"alpha", "beta"};
    const std::string rep = t.report(names, 2);
    expect(rep.find("0 alpha 20 1") != std::string::npos && rep.find("0 beta 5 1") != std::string::npos,
           "report names the settled choices");

    // a tuning cache: the settled choices of one tuner taken by another, which
    // then runs them without trials; unsettled cells and unknown knobs skipped
    pra::Tuner u(3);
    const std::string cache = t.report(names, 2) + "0 alpha 30 -1 1.0\n0 gamma 20 1 1.0\n";
    expect(u.load(cache, names, 2) == 2, "two settled choices read back");
    expect(u.settled(0, 0, 20) == 1 && u.settled(0, 1, 5) == 1, "read-back choices are settled");
    expect(u.settled(0, 0, 30) == -1, "an unsettled cell is not taken");
    expect(u.pick(0, 0, 20, 3) == 1, "a read-back choice runs without trials");

    if (failures) return 1;
    std::printf("ok\n");
    return 0;
}
