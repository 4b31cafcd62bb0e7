// hp2d_dem.cpp - depression filling of a DEM on the host (preprocessing).
//
// Priority-Flood+epsilon (R. Barnes, C. Lehman, D. Mulla, "Priority-Flood: An
// optimal depression-filling and watershed-labeling algorithm for digital
// elevation models", Computers & Geosciences 62, 2014, Algorithm 3): cells are
// popped in order of elevation from a priority queue seeded with the drainage
// cells; each unvisited neighbour (D8) is raised to at least the popped
// elevation plus a small increment, so every cell keeps a strictly descending
// path to a seed and the routing sees no flats.  NaN cells (outside the
// domain) are left NaN and never entered.
#include <cmath>
#include <cstdint>
#include <algorithm>
#include <functional>
#include <numeric>
#include <queue>
#include <vector>

#include "hp2d.h"

extern "C" hp2d_status hp2d_fill_sinks(int32_t ny, int32_t nx, const double *z,
                                       const uint8_t *seed_mask, double epsilon, double *out) {
    if (ny < 1 || nx < 1 || !z || !out || !(epsilon >= 0.0)) return HP2D_E_ARG;
    const long long n = (long long)ny * nx;
    std::vector<uint8_t> closed(n, 0);
    for (long long c = 0; c < n; ++c) {
        out[c] = z[c];
        if (std::isnan(z[c])) closed[c] = 1;
    }
    using Item = std::pair<double, long long>;  // (elevation, cell), smallest first
    std::priority_queue<Item, std::vector<Item>, std::greater<Item>> open;
    auto seed = [&](long long c) {
        if (!closed[c]) {
            closed[c] = 1;
            open.push({out[c], c});
        }
    };
    if (seed_mask) {
        for (long long c = 0; c < n; ++c)
            if (seed_mask[c]) seed(c);
    } else {
        for (int j = 0; j < nx; ++j) {
            seed((long long)j * ny);
            seed((long long)j * ny + ny - 1);
        }
        for (int i = 0; i < ny; ++i) {
            seed(i);
            seed((long long)(nx - 1) * ny + i);
        }
    }
    // cells next to the domain mask drain as well when the mask is irregular
    if (!seed_mask)
        for (long long c = 0; c < n; ++c) {
            if (closed[c]) continue;
            const int i = (int)(c % ny), j = (int)(c / ny);
            bool edge = false;
            for (int di = -1; di <= 1 && !edge; ++di)
                for (int dj = -1; dj <= 1; ++dj) {
                    const int a = i + di, b = j + dj;
                    if (a < 0 || a >= ny || b < 0 || b >= nx) continue;
                    if (std::isnan(z[(long long)b * ny + a])) edge = true;
                }
            if (edge) seed(c);
        }
    while (!open.empty()) {
        const Item it = open.top();
        open.pop();
        const long long c = it.second;
        const int i = (int)(c % ny), j = (int)(c / ny);
        for (int di = -1; di <= 1; ++di)
            for (int dj = -1; dj <= 1; ++dj) {
                if (!di && !dj) continue;
                const int a = i + di, b = j + dj;
                if (a < 0 || a >= ny || b < 0 || b >= nx) continue;
                const long long q = (long long)b * ny + a;
                if (closed[q]) continue;
                closed[q] = 1;
                const double lift = std::nextafter(out[c] + epsilon, INFINITY);
                if (out[q] <= out[c]) out[q] = epsilon > 0.0 ? lift : out[c];
                open.push({out[q], q});
            }
    }
    return HP2D_OK;
}

extern "C" hp2d_status hp2d_flow_accumulation(int32_t ny, int32_t nx, const double *z,
                                              double *out) {
    if (ny < 1 || nx < 1 || !z || !out) return HP2D_E_ARG;
    const long long n = (long long)ny * nx;
    // D8 receiver of every valid cell: the steepest descent (drop / distance,
    // diagonals at sqrt(2)); -1 where no neighbour is lower
    std::vector<long long> recv((size_t)n, -1);
    std::vector<int> donors((size_t)n, 0);
    const double diag = std::sqrt(2.0);
    for (long long c = 0; c < n; ++c) {
        out[c] = std::isnan(z[c]) ? 0.0 : 1.0;
        if (std::isnan(z[c])) continue;
        const int i = (int)(c % ny), j = (int)(c / ny);
        double best = 0.0;
        long long to = -1;
        for (int dj = -1; dj <= 1; ++dj)
            for (int di = -1; di <= 1; ++di) {
                if (!di && !dj) continue;
                const int a = i + di, b = j + dj;
                if (a < 0 || a >= ny || b < 0 || b >= nx) continue;
                const long long q = (long long)b * ny + a;
                if (std::isnan(z[q])) continue;
                const double s = (z[c] - z[q]) / ((di && dj) ? diag : 1.0);
                if (s > best) {
                    best = s;
                    to = q;
                }
            }
        recv[(size_t)c] = to;
        if (to >= 0) donors[(size_t)to] += 1;
    }
    // pass counts downstream in topological order (a cell goes once all its
    // donors have); counts are integers, so the order does not change the sums
    std::vector<long long> ready;
    ready.reserve((size_t)n);
    for (long long c = 0; c < n; ++c)
        if (!std::isnan(z[c]) && donors[(size_t)c] == 0) ready.push_back(c);
    for (size_t k = 0; k < ready.size(); ++k) {
        const long long c = ready[k];
        const long long to = recv[(size_t)c];
        if (to < 0) continue;
        out[to] += out[c];
        if (--donors[(size_t)to] == 0) ready.push_back(to);
    }
    return HP2D_OK;
}
