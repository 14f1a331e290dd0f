// Runs the system libstdc++'s real std::sort the way SMAX::Utils::
// convert_crs_to_scs does (src/utils/convert_formats.hpp:10-13,46-55): on
// {int index; int count;} pairs, per sigma window, comparator
// a.count > b.count.  Used by the tests to pin the oracle's restatement of
// that sort and to write the tie-order digests in sort_digests.json.
//
//   sort_pin sort <sigma> <counts.bin> <index_out.bin>
//       counts.bin: int32 row lengths (already padded to n_rows_padded)
//   sort_pin killer <n> <counts_out.bin>
//       McIlroy's adversary ("A Killer Adversary for Quicksort", 1999) played
//       against std::sort with the same comparator: the returned counts drive
//       introsort to its depth limit, so its heap-sort fallback runs.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

struct Pair {
    int index;
    int count;
};

static std::vector<int> read_i32(const char *path) {
    FILE *f = std::fopen(path, "rb");
    if (!f) { std::perror(path); std::exit(2); }
    std::fseek(f, 0, SEEK_END);
    long bytes = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    std::vector<int> v(static_cast<size_t>(bytes) / 4);
    if (!v.empty() && std::fread(v.data(), 4, v.size(), f) != v.size()) {
        std::perror("read");
        std::exit(2);
    }
    std::fclose(f);
    return v;
}

static void write_i32(const char *path, const std::vector<int> &v) {
    FILE *f = std::fopen(path, "wb");
    if (!f) { std::perror(path); std::exit(2); }
    if (!v.empty()) std::fwrite(v.data(), 4, v.size(), f);
    std::fclose(f);
}

// ---------------------------------------------------------------- adversary
static std::vector<int> g_val;
static int g_gas, g_nsolid, g_candidate;

static int adversary_cmp(int x, int y) {
    if (g_val[x] == g_gas && g_val[y] == g_gas) {
        if (x == g_candidate) g_val[x] = g_nsolid++;
        else g_val[y] = g_nsolid++;
    }
    if (g_val[x] == g_gas) g_candidate = x;
    else if (g_val[y] == g_gas) g_candidate = y;
    return g_val[x] - g_val[y];
}

int main(int argc, char **argv) {
    if (argc >= 5 && !std::strcmp(argv[1], "sort")) {
        long sigma = std::atol(argv[2]);
        std::vector<int> cnt = read_i32(argv[3]);
        const long n = static_cast<long>(cnt.size());
        std::vector<Pair> e(cnt.size());
        for (long i = 0; i < n; ++i) e[i] = Pair{static_cast<int>(i), cnt[i]};
        for (long i = 0; i < n; i += sigma) {
            auto begin = e.begin() + i;
            auto end = (i + sigma) < n ? e.begin() + i + sigma : e.end();
            std::sort(begin, end,
                      [](const Pair &a, const Pair &b) { return a.count > b.count; });
        }
        std::vector<int> idx(cnt.size());
        for (long i = 0; i < n; ++i) idx[i] = e[i].index;
        write_i32(argv[4], idx);
        return 0;
    }
    if (argc >= 4 && !std::strcmp(argv[1], "killer")) {
        const int n = std::atoi(argv[2]);
        g_val.assign(n, n - 1);
        g_gas = n - 1;
        g_nsolid = 0;
        g_candidate = 0;
        std::vector<int> ptr(n);
        for (int i = 0; i < n; ++i) ptr[i] = i;
        // count = n-1-val: "a.count > b.count" answers "val[a] < val[b]".
        std::sort(ptr.begin(), ptr.end(),
                  [](int a, int b) { return adversary_cmp(a, b) < 0; });
        std::vector<int> cnt(n);
        for (int i = 0; i < n; ++i) cnt[i] = n - 1 - g_val[i];
        write_i32(argv[3], cnt);
        return 0;
    }
    std::fprintf(stderr, "usage: sort_pin sort <sigma> <in> <out> | killer <n> <out>\n");
    return 2;
}
