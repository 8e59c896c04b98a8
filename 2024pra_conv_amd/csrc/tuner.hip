// Tuned tile-configuration table consulted by getkernelInfo.
//
// Replaces the reference's hand-written (n, c) if/else dispatch (src/conv2d.cpp:273-494):
// a table of exact shapes -> GEMM tile configuration, filled by the autotuner (or loaded
// from its saved file). Shapes not in the table fall back to the shape heuristic.
#include <mutex>
#include <stdio.h>
#include <string.h>
#include <vector>

#include "plugin_internal.h"
#include "wino_gemm.h"

namespace {

struct TunedEntry {
    unsigned n, c, h, w, k, p, q;
    int cfg;
};

std::mutex g_mu;
std::vector<TunedEntry> g_table;

bool same_shape(const TunedEntry& e, const problem_t* pb) {
    return e.n == pb->n && e.c == pb->c && e.h == pb->h && e.w == pb->w && e.k == pb->k &&
           e.p == pb->p && e.q == pb->q;
}

}  // namespace

namespace wino {

int tuned_config_lookup(const problem_t* pb) {
    std::lock_guard<std::mutex> lk(g_mu);
    for (const auto& e : g_table)
        if (same_shape(e, pb)) return e.cfg;
    return -1;
}

}  // namespace wino

extern "C" {

// Records (or replaces) the configuration for one exact shape.
int conv2d_tuned_set(const problem_t* pb, int cfg) {
    if (!pb) return CONV2D_ERR_NULL;
    if (cfg < 0 || cfg >= kNumGemmConfigs) return CONV2D_ERR_INVALID;
    std::lock_guard<std::mutex> lk(g_mu);
    for (auto& e : g_table)
        if (same_shape(e, pb)) {
            e.cfg = cfg;
            return CONV2D_OK;
        }
    g_table.push_back({pb->n, pb->c, pb->h, pb->w, pb->k, pb->p, pb->q, cfg});
    return CONV2D_OK;
}

// Empties the table (every shape falls back to the heuristic).
int conv2d_tuned_clear(void) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_table.clear();
    return CONV2D_OK;
}

int conv2d_tuned_count(void) {
    std::lock_guard<std::mutex> lk(g_mu);
    return (int)g_table.size();
}

// Text format, one shape per line: "n c h w k p q cfg". Lines starting with '#' are
// comments.
int conv2d_tuned_save(const char* path) {
    if (!path) return CONV2D_ERR_NULL;
    FILE* f = fopen(path, "w");
    if (!f) return CONV2D_ERR_INVALID;
    std::lock_guard<std::mutex> lk(g_mu);
    fprintf(f, "# winoconv tuned tile configurations: n c h w k p q cfg\n");
    for (const auto& e : g_table)
        fprintf(f, "%u %u %u %u %u %u %u %d\n", e.n, e.c, e.h, e.w, e.k, e.p, e.q, e.cfg);
    fclose(f);
    return CONV2D_OK;
}

int conv2d_tuned_load(const char* path) {
    if (!path) return CONV2D_ERR_NULL;
    FILE* f = fopen(path, "r");
    if (!f) return CONV2D_ERR_INVALID;
    char line[256];
    std::vector<TunedEntry> rows;
    int rc = CONV2D_OK;
    while (fgets(line, sizeof line, f)) {
        if (line[0] == '#' || line[0] == '\n') continue;
        TunedEntry e;
        if (sscanf(line, "%u %u %u %u %u %u %u %d", &e.n, &e.c, &e.h, &e.w, &e.k, &e.p, &e.q,
                   &e.cfg) != 8 ||
            e.cfg < 0 || e.cfg >= kNumGemmConfigs) {
            rc = CONV2D_ERR_INVALID;
            break;
        }
        rows.push_back(e);
    }
    fclose(f);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(g_mu);
    for (const auto& r : rows) {
        bool found = false;
        for (auto& e : g_table)
            if (e.n == r.n && e.c == r.c && e.h == r.h && e.w == r.w && e.k == r.k &&
                e.p == r.p && e.q == r.q) {
                e.cfg = r.cfg;
                found = true;
            }
        if (!found) g_table.push_back(r);
    }
    return CONV2D_OK;
}

}  // extern "C"
