// Launch-shape table of the engine, written by the project's autotuner
// (paco-svd_amd/autotune.py sweeps the candidates on the GPU and writes
// autotune_table.json next to the library; the table is read once at first use).
// Keys are size buckets: M and N rounded up to powers of two.  Values: threads of the
// one-workgroup bidiagonalization for matrices too large to sit in LDS whole.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <dlfcn.h>
#include <map>
#include <mutex>
#include <string>
#include <utility>

namespace paco {

static std::map<std::pair<int, int>, int> g_table;
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
    // entries of the form  "MxN": threads  (a flat JSON object)
    char key[64];
    int m, n, t;
    int c;
    while ((c = fgetc(f)) != EOF) {
        if (c != '"') continue;
        int len = 0;
        while ((c = fgetc(f)) != EOF && c != '"' && len < 63) key[len++] = (char)c;
        key[len] = 0;
        if (sscanf(key, "%dx%d", &m, &n) != 2) continue;
        if (fscanf(f, " : %d", &t) == 1 && (t == 256 || t == 512 || t == 1024)) g_table[{m, n}] = t;
    }
    fclose(f);
}

int tuned_bidiag_threads(int M, int N)
{
    std::call_once(g_once, load_table);
    if (const char *env = getenv("PACO_BIDIAG_THREADS")) {
        const int t = atoi(env);
        if (t == 256 || t == 512 || t == 1024) return t;
    }
    auto it = g_table.find({bucket(M), bucket(N)});
    if (it != g_table.end()) return it->second;
    return 1024;
}

} // namespace paco
