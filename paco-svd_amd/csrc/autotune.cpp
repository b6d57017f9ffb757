// Launch-shape table of the engine, written by the project's autotuner
// (paco-svd_amd/autotune.py sweeps the candidates on the GPU and writes
// autotune_table.json next to the library; the table is read once at first use).
// A flat JSON object of integer entries; size buckets round M and N up to powers of two:
//   "MxN": t              threads of the one-workgroup bidiagonalization of a streamed
//                         (too large for LDS) matrix of the bucket (256, 512, 1024)
//   "res_threads:c": t    threads of the LDS-resident launches of LDS class c (0..3)
//   "leaf_size": ls       columns per D&C leaf (3..32), the library's default
//   "grid_tiles:MxN": g   tiles per side of the chip-wide bidiagonalization (4..64)
//   "grid_switch:MxN": s  columns left to the one-workgroup kernel by the chip-wide path
// The PACO_* environment variables of each parameter take precedence over the table.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <dlfcn.h>
#include <map>
#include <mutex>
#include <string>
#include <utility>

namespace paco {

struct TunedTable {
    std::map<std::pair<int, int>, int> threads, grid_tiles, grid_switch;
    int res_threads[4] = {0, 0, 0, 0};
    int leaf_size = 0;
};
static TunedTable g_table;
static std::once_flag g_once;

static int bucket(int x)
{
    int b = 1;
    while (b < x) b <<= 1;
    return b;
}

static void load_table()
{
    std::string path;
    if (const char *env = getenv("PACO_AUTOTUNE_TABLE")) {
        path = env;
    } else {
        Dl_info info;
        if (dladdr((void *)&load_table, &info) && info.dli_fname) {
            path = info.dli_fname;
            const size_t k = path.find_last_of('/');
            path = (k == std::string::npos ? std::string(".") : path.substr(0, k)) + "/autotune_table.json";
        }
    }
    FILE *f = path.empty() ? nullptr : fopen(path.c_str(), "r");
    if (!f) return;
    char key[64];
    int m, n, t;
    int c;
    while ((c = fgetc(f)) != EOF) {
        if (c != '"') continue;
        int len = 0;
        while ((c = fgetc(f)) != EOF && c != '"' && len < 63) key[len++] = (char)c;
        key[len] = 0;
        if (fscanf(f, " : %d", &t) != 1) continue;
        if (sscanf(key, "grid_tiles:%dx%d", &m, &n) == 2) {
            if (t >= 4 && t <= 64) g_table.grid_tiles[{m, n}] = t;
        } else if (sscanf(key, "grid_switch:%dx%d", &m, &n) == 2) {
            if (t >= 32 && t <= 512) g_table.grid_switch[{m, n}] = t;
        } else if (sscanf(key, "res_threads:%d", &n) == 1) {
            if (n >= 0 && n < 4 && (t == 256 || t == 512 || t == 1024)) g_table.res_threads[n] = t;
        } else if (!strcmp(key, "leaf_size")) {
            if (t >= 3 && t <= 32) g_table.leaf_size = t;
        } else if (sscanf(key, "%dx%d", &m, &n) == 2) {
            if (t == 256 || t == 512 || t == 1024) g_table.threads[{m, n}] = t;
        }
    }
    fclose(f);
}

static const TunedTable &table()
{
    std::call_once(g_once, load_table);
    return g_table;
}

int tuned_bidiag_threads(int M, int N)
{
    if (const char *env = getenv("PACO_BIDIAG_THREADS")) {
        const int t = atoi(env);
        if (t == 256 || t == 512 || t == 1024) return t;
    }
    const auto &T = table().threads;
    auto it = T.find({bucket(M), bucket(N)});
    return it != T.end() ? it->second : 1024;
}

int tuned_res_threads(int cls, int fallback)
{
    const int t = (cls >= 0 && cls < 4) ? table().res_threads[cls] : 0;
    return t ? t : fallback;
}

int tuned_leaf_size(int fallback)
{
    const int t = table().leaf_size;
    return t ? t : fallback;
}

int tuned_grid_tiles(int M, int N, int fallback)
{
    const auto &T = table().grid_tiles;
    auto it = T.find({bucket(M), bucket(N)});
    return it != T.end() ? it->second : fallback;
}

int tuned_grid_switch(int M, int N, int fallback)
{
    const auto &T = table().grid_switch;
    auto it = T.find({bucket(M), bucket(N)});
    return it != T.end() ? it->second : fallback;
}

} // namespace paco
