// Launch-shape table written by the autotuner (autotune.py) and read here at run time.
//
// The table is text, one entry per line:
//     <kind> <nodes> <particles> key=value key=value ...
// ('#' starts a comment).  kind names what the entry tunes: "push" (particles per push
// workgroup, map entries per XCD stripe), "mg" (smoothing sweep pairs, first Galerkin level,
// largest level of the single-workgroup coarse cycle), "ensemble" (particles per push
// workgroup, LDS deposit window, SOR workgroup size, steps between sorts).  A lookup takes the
// entry of that kind nearest to the request in (log nodes, log particles); without a table, or
// without an entry of that kind, the caller's default applies.
//
// Location: the file named by PICAMD_TUNING, else tuning/table.txt beside libpicamd.so.
#include <dlfcn.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "picamd_internal.h"

namespace pic {

struct TuneEntry {
    std::string kind;
    double nodes = 0, particles = 0;
    std::map<std::string, double> v;
};

static std::mutex g_tune_mu;
static bool g_tune_loaded = false;
static std::vector<TuneEntry> g_tune;
static std::string g_tune_path;

static std::string default_table_path() {
    const char *env = getenv("PICAMD_TUNING");
    if (env && env[0]) return env;
    Dl_info info;
    if (dladdr((void *)&default_table_path, &info) && info.dli_fname) {
        std::string so = info.dli_fname;
        size_t slash = so.find_last_of('/');
        std::string dir = slash == std::string::npos ? "." : so.substr(0, slash);
        return dir + "/tuning/table.txt";
    }
    return "tuning/table.txt";
}

// parses the table at path into `out`; returns false when the file cannot be read
static bool parse_table(const std::string &path, std::vector<TuneEntry> &out) {
    FILE *f = fopen(path.c_str(), "r");
    if (!f) return false;
    char line[4096];
    while (fgets(line, sizeof line, f)) {
        char *hash = strchr(line, '#');
        if (hash) *hash = 0;
        std::istringstream in(line);
        TuneEntry e;
        if (!(in >> e.kind >> e.nodes >> e.particles)) continue;
        std::string kv;
        while (in >> kv) {
            size_t eq = kv.find('=');
            if (eq == std::string::npos) continue;
            e.v[kv.substr(0, eq)] = atof(kv.c_str() + eq + 1);
        }
        out.push_back(e);
    }
    fclose(f);
    return true;
}

static void ensure_loaded() {
    if (g_tune_loaded) return;
    g_tune_loaded = true;
    g_tune_path = default_table_path();
    parse_table(g_tune_path, g_tune);
}

double tuned(const char *kind, double nodes, double particles, const char *key, double dflt) {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    ensure_loaded();
    const TuneEntry *best = nullptr;
    double bd = 0;
    auto lg = [](double x) { return std::log(std::max(x, 1.0)); };
    for (const TuneEntry &e : g_tune) {
        if (e.kind != kind || !e.v.count(key)) continue;
        double dn = lg(e.nodes) - lg(nodes), dp = lg(e.particles) - lg(particles);
        double d = dn * dn + dp * dp;
        if (!best || d < bd) {
            best = &e;
            bd = d;
        }
    }
    return best ? best->v.at(key) : dflt;
}

}  // namespace pic

using namespace pic;

extern "C" {

int pic_tuning_load(const char *path) {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    std::string p = path && path[0] ? std::string(path) : default_table_path();
    std::vector<TuneEntry> t;
    if (!parse_table(p, t)) {
        set_error("tuning table not readable: " + p);
        return PIC_ERR_ARG;
    }
    g_tune.swap(t);
    g_tune_path = p;
    g_tune_loaded = true;
    return PIC_OK;
}

int pic_tuning_lookup(const char *kind, double nodes, double particles, const char *key, double dflt,
                      double *value) {
    PIC_CHECK(kind && key && value, PIC_ERR_ARG, "null argument");
    *value = tuned(kind, nodes, particles, key, dflt);
    return PIC_OK;
}

}  // extern "C"
