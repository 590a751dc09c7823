// Autotuner state, kept free of HIP types so that its rule can be exercised
// on the CPU.
//
// A knob is a choice between variants of one step that give identical
// results (a table's load factor, a kernel's block width) and differ only in
// speed.  For each (device, knob, size band) the tuner keeps the mean
// measured cost per unit of work of every variant: until every variant has
// `samples` measurements, pick() returns the variant with the fewest (lowest
// index first), so the variants take turns on live iterations; after that it
// returns the cheapest, for good.  Bands let one knob settle differently for
// small and large work.
#ifndef PRA_TUNE_HPP
#define PRA_TUNE_HPP

#include <cstdio>
#include <map>
#include <mutex>
#include <string>
#include <tuple>

namespace pra {

class Tuner {
  public:
    static constexpr int kMaxVariants = 4;

    explicit Tuner(int samples = 3) : samples_(samples) {}

    // the variant to run next for (device, knob, band) among nvar variants
    int pick(int device, int knob, int band, int nvar)
    {
        std::lock_guard<std::mutex> lk(mu_);
        Cell& c = cells_[{device, knob, band}];
        if (c.choice >= 0) return c.choice;
        int best = 0;
        for (int v = 1; v < nvar; v++)
            if (c.n[v] < c.n[best]) best = v;
        return best;
    }

    // one measurement of variant var: `cost` per unit of work; true when it
    // settled the cell
    bool record(int device, int knob, int band, int var, double cost, int nvar)
    {
        if (var < 0 || var >= nvar || nvar > kMaxVariants || !(cost >= 0)) return false;
        std::lock_guard<std::mutex> lk(mu_);
        Cell& c = cells_[{device, knob, band}];
        if (c.choice >= 0) return false;
        c.sum[var] += cost;
        c.n[var]++;
        for (int v = 0; v < nvar; v++)
            if (c.n[v] < samples_) return false;
        int best = 0;
        for (int v = 1; v < nvar; v++)
            if (c.sum[v] / c.n[v] < c.sum[best] / c.n[best]) best = v;
        c.choice = best;
        return true;
    }

    // a choice settled earlier (a tuning cache read back): no more trials
    void set(int device, int knob, int band, int choice)
    {
        if (choice < 0 || choice >= kMaxVariants) return;
        std::lock_guard<std::mutex> lk(mu_);
        cells_[{device, knob, band}].choice = choice;
    }

    // the settled choices of a report ("device knob band choice ..." lines;
    // unknown knobs and unsettled cells are skipped); returns the count taken
    int load(const std::string& text, const char* const* knob_names, int nknobs)
    {
        int taken = 0;
        size_t at = 0;
        while (at < text.size()) {
            size_t end = text.find('\n', at);
            if (end == std::string::npos) end = text.size();
            const std::string line = text.substr(at, end - at);
            at = end + 1;
            int device = 0, band = 0, choice = -1;
            char name[96] = {0};
            if (std::sscanf(line.c_str(), "%d %95s %d %d", &device, name, &band, &choice) != 4 || choice < 0) continue;
            for (int k = 0; k < nknobs; k++)
                if (std::string(name) == knob_names[k]) {
                    set(device, k, band, choice);
                    taken++;
                }
        }
        return taken;
    }

    // the settled variant, or -1 while the variants are still being tried
    int settled(int device, int knob, int band) const
    {
        std::lock_guard<std::mutex> lk(mu_);
        auto it = cells_.find({device, knob, band});
        return it == cells_.end() ? -1 : it->second.choice;
    }

    // one line per (device, knob, band): "device knob band choice mean0 mean1 ..."
    std::string report(const char* const* knob_names, int nknobs) const
    {
        std::lock_guard<std::mutex> lk(mu_);
        std::string out;
        for (const auto& e : cells_) {
            const int knob = std::get<1>(e.first);
            char line[256];
            int k = std::snprintf(line, sizeof(line), "%d %s %d %d", std::get<0>(e.first),
                                  knob >= 0 && knob < nknobs ? knob_names[knob] : "?", std::get<2>(e.first),
                                  e.second.choice);
            for (int v = 0; v < kMaxVariants && k > 0 && k < (int)sizeof(line); v++)
                if (e.second.n[v])
                    k += std::snprintf(line + k, sizeof(line) - (size_t)k, " %.4g", e.second.sum[v] / e.second.n[v]);
            out += line;
            out += "\n";
        }
        return out;
    }

  private:
    struct Cell {
        double sum[kMaxVariants] = {0, 0, 0, 0};
        int n[kMaxVariants] = {0, 0, 0, 0};
        int choice = -1;
    };
    int samples_;
    mutable std::mutex mu_;
    std::map<std::tuple<int, int, int>, Cell> cells_;
};

} // namespace pra

#endif
