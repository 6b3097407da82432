// `blest` command line: the reference's main/Benchmark flow (main.cu:19-131,
// Benchmark.cuh:27-331) over the C ABI of libblest.so only.
//
//   blest -d <directory> -g <graph> -j <0|1> -w <window> [options]
//
// <graph> names a local MatrixMarket file <directory>/<graph>.mtx (nothing is
// downloaded) or a synthetic graph: rmat:<scale>[:<edge factor>[:<seed>]],
// grid:<side>[:<keep>[:<shortcut fraction>[:<seed>]]], er:<n>[:<avg degree>[:<seed>]].
// Sources come from <directory>/<graph>.txt (one original vertex ID per line,
// Benchmark.cuh:67-85) or, with --roots K, K seeded roots of degree >= 1.
// Options: --kernel BFS|CC|MBFS (default BFS), --gorder (social graphs: Gorder+Jaccard),
//          --undirected 0|1 (MTX, default 1), --no-cache, --roots K, --seed S.
// Caches (Benchmark.cuh:163-209): <directory>/<graph>_csc.bin, _ordering.bin; the
// BVSS statistics and one row per source go to <directory>/<graph>.csv.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <limits>
#include <sstream>
#include <stdexcept>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "blest.h"

namespace {

void usage(const char* prog) {
    std::cerr << "Usage: " << prog << " -d <directory> -g <graph_name> -j <0|1> -w <window_size> [options]\n"
              << "  -d \t Directory holding <graph>.mtx / <graph>.txt; caches and <graph>.csv are written there\n"
              << "  -g \t Graph: a local MatrixMarket name, or rmat:<scale>[:ef[:seed]], grid:<side>[:keep[:frac[:seed]]],\n"
              << "     \t er:<n>[:deg[:seed]]\n"
              << "  -j \t Jaccard enabled (0 or 1)\n"
              << "  -w \t Window size (an unsigned integer)\n"
              << "  --kernel BFS|CC|MBFS   --gorder   --roots K   --seed S   --undirected 0|1   --no-cache\n";
}

struct Config {
    std::string directory, graph, kernel = "BFS";
    bool jaccard = false, jaccardSet = false, windowSet = false, gorder = false, cache = true, undirected = true;
    unsigned window = 0, roots = 0;
    unsigned long long seed = 1;
};

Config parse(int argc, char** argv) {
    Config c;
    auto value = [&](int& i, const char* opt) -> std::string {
        if (i + 1 >= argc) throw std::runtime_error(std::string("Missing value for ") + opt);
        return argv[++i];
    };
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "-h" || a == "--help") {
            usage(argv[0]);
            std::exit(0);
        } else if (a == "-d") {
            c.directory = value(i, "-d");
        } else if (a == "-g") {
            c.graph = value(i, "-g");
        } else if (a == "-j") {
            const std::string v = value(i, "-j");
            if (v != "0" && v != "1") throw std::runtime_error("Invalid value for -j (expected 0 or 1), got: " + v);
            c.jaccard = v == "1";
            c.jaccardSet = true;
        } else if (a == "-w") {
            const std::string v = value(i, "-w");
            try {
                size_t pos = 0;
                unsigned long w = std::stoul(v, &pos);
                if (pos != v.size() || w > std::numeric_limits<unsigned>::max()) throw std::out_of_range("w");
                c.window = unsigned(w);
            } catch (const std::exception&) {
                throw std::runtime_error("Invalid value for -w (expected unsigned integer), got: " + v);
            }
            c.windowSet = true;
        } else if (a == "--kernel") {
            c.kernel = value(i, "--kernel");
            if (c.kernel != "BFS" && c.kernel != "CC" && c.kernel != "MBFS")
                throw std::runtime_error("Invalid --kernel (BFS, CC or MBFS): " + c.kernel);
        } else if (a == "--gorder") {
            c.gorder = true;
        } else if (a == "--no-cache") {
            c.cache = false;
        } else if (a == "--roots") {
            c.roots = unsigned(std::stoul(value(i, "--roots")));
        } else if (a == "--seed") {
            c.seed = std::stoull(value(i, "--seed"));
        } else if (a == "--undirected") {
            c.undirected = value(i, "--undirected") != "0";
        } else {
            throw std::runtime_error("Unknown argument: " + a);
        }
    }
    if (c.directory.empty()) throw std::runtime_error("Missing required option: -d <directory>");
    if (c.graph.empty()) throw std::runtime_error("Missing required option: -g <graph_name>");
    if (!c.jaccardSet) throw std::runtime_error("Missing required option: -j <0|1>");
    if (!c.windowSet) throw std::runtime_error("Missing required option: -w <window_size>");
    if (!c.directory.empty() && c.directory.back() != '/') c.directory += '/';
    return c;
}

void check(blest_status s) {
    if (s != BLEST_OK) throw std::runtime_error(std::string(blest_status_str(s)) + ": " + blest_last_error());
}

bool exists(const std::string& p) {
    struct stat st;
    return ::stat(p.c_str(), &st) == 0;
}

std::vector<std::string> split(const std::string& s, char sep) {
    std::vector<std::string> out;
    std::stringstream ss(s);
    std::string x;
    while (std::getline(ss, x, sep)) out.push_back(x);
    return out;
}

// Synthetic graph spec or a local MatrixMarket file.
blest_csc* load_graph(const Config& c, const std::string& base, bool& synthetic) {
    blest_csc* g = nullptr;
    const auto parts = split(c.graph, ':');
    synthetic = true;
    if (parts[0] == "rmat" && parts.size() >= 2) {
        unsigned scale = unsigned(std::stoul(parts[1]));
        unsigned ef = parts.size() > 2 ? unsigned(std::stoul(parts[2])) : 16;
        unsigned long long seed = parts.size() > 3 ? std::stoull(parts[3]) : 1;
        check(blest_csc_generate_rmat(scale, ef, 0.57, 0.19, 0.19, seed, 1, &g));
    } else if (parts[0] == "grid" && parts.size() >= 2) {
        unsigned side = unsigned(std::stoul(parts[1]));
        double keep = parts.size() > 2 ? std::stod(parts[2]) : 0.9;
        double frac = parts.size() > 3 ? std::stod(parts[3]) : 0.01;
        unsigned long long seed = parts.size() > 4 ? std::stoull(parts[4]) : 1;
        check(blest_csc_generate_grid(side, keep, frac, seed, &g));
    } else if (parts[0] == "er" && parts.size() >= 2) {
        unsigned n = unsigned(std::stoul(parts[1]));
        double deg = parts.size() > 2 ? std::stod(parts[2]) : 8.0;
        unsigned long long seed = parts.size() > 3 ? std::stoull(parts[3]) : 1;
        check(blest_csc_generate_er(n, deg, seed, &g));
    } else {
        synthetic = false;
        const std::string cache = base + "_csc.bin";
        if (c.cache && exists(cache)) {
            check(blest_csc_load_binary(cache.c_str(), &g));
            std::cout << "CSC read from binary." << std::endl;
        } else {
            std::cout << "CSC construction started." << std::endl;
            check(blest_csc_from_mtx((base + ".mtx").c_str(), c.undirected, 1, &g));
            if (c.cache) check(blest_csc_save_binary(g, cache.c_str()));
            std::cout << "CSC constructed." << std::endl;
        }
    }
    return g;
}

int run(const Config& c) {
    std::string name = c.graph;
    for (char& ch : name)
        if (ch == ':' || ch == '/') ch = '_';
    const std::string base = c.directory + name;
    bool synthetic = false;
    blest_csc* g = load_graph(c, c.directory + c.graph, synthetic);
    uint32_t n = 0;
    uint64_t nnz = 0;
    int social = 0;
    check(blest_csc_info(g, &n, &nnz, &social));
    std::cout << "Number of vertices: " << n << std::endl << "Number of edges: " << nnz << std::endl;
    std::cout << (social ? "The graph is a social network." : "The graph is not a social network.") << std::endl;

    // Ordering: cached inverse permutation or CSC::reorder (Benchmark.cuh:196-209).
    std::vector<uint32_t> inv(n);
    const std::string ordCache = base + "_ordering.bin";
    if (c.cache && exists(ordCache)) {
        check(blest_ordering_load(ordCache.c_str(), inv.data(), n));
        check(blest_csc_apply_permutation(g, inv.data()));
        std::cout << "Ordering read from binary." << std::endl;
    } else {
        blest_reorder_opts o;
        blest_reorder_opts_default(&o);
        o.kind = BLEST_REORDER_AUTO;
        o.slice_size = 8;
        o.window = c.window;
        o.jaccard = c.jaccard ? 1 : 0;
        o.gorder = c.gorder ? 1 : 0;
        o.seed = c.seed;
        check(blest_csc_reorder(g, &o, inv.data()));
        if (c.cache) check(blest_ordering_save(ordCache.c_str(), inv.data(), n));
    }

    // BVSS (Benchmark.cuh:212-228), FULL_PADDING = !social (Benchmark.cuh:187-194).
    blest_bvss* b = nullptr;
    check(blest_bvss_build(g, 8, 4, social ? 0 : 1, &b));
    blest_bvss_info bi;
    check(blest_bvss_info_get(b, &bi));
    std::cout << "Number of real slice sets: " << bi.no_real_slice_sets << std::endl
              << "Number of virtual slice sets: " << bi.no_virtual_slice_sets << std::endl
              << "Number of slices: " << bi.no_slices << std::endl
              << "Update divergence is: " << bi.update_divergence << std::endl
              << "Check: " << bi.edges_check << std::endl;
    std::ofstream csv(base + ".csv");
    char buf[4096];
    check(blest_bvss_csv_header(buf, sizeof buf));
    csv << buf << std::endl;
    check(blest_bvss_csv_row(b, buf, sizeof buf));
    csv << buf << std::endl << "Source\tTotalLevels\tTotalVisited\tTime(ms)" << std::endl;

    blest_device_graph* dg = nullptr;
    blest_device_opts dopt;
    blest_device_opts_default(&dopt);
    check(blest_device_graph_create(b, g, inv.data(), &dopt, &dg));

    // Sources: the reference's file (original IDs) or seeded roots.
    std::vector<uint32_t> sources;
    if (c.roots) {
        // roots are drawn on the reordered graph; report them in original IDs
        std::vector<uint32_t> r(c.roots), oldOf(n);
        check(blest_csc_pick_roots(g, c.roots, c.seed, r.data()));
        for (uint32_t old = 0; old < n; ++old) oldOf[inv[old]] = old;
        for (uint32_t v : r) sources.push_back(oldOf[v]);
    } else {
        std::ifstream f(c.directory + c.graph + ".txt");
        if (!f) throw std::runtime_error("Failed to open file from which to read source vertices.");
        uint64_t s;
        while (f >> s) {
            if (s >= n) throw std::runtime_error("source vertex out of range: " + std::to_string(s));
            sources.push_back(uint32_t(s));
        }
    }

    double total = 0;
    int rc = 0;
    blest_query* q = nullptr;
    if (c.kernel == "BFS") {
        std::cout << "BFS kernels launching..." << std::endl;
        check(blest_query_create(dg, 1, &q));
        std::vector<int32_t> levels(n);
        for (uint32_t s : sources) {
            blest_bfs_stats st;
            check(blest_bfs(q, s, levels.data(), &st));
            std::cout << "Source: " << s << " - Number of levels processed: " << st.total_levels
                      << " - Total visited: " << st.no_visited << " - Time took: " << st.kernel_ms << " ms."
                      << std::endl;
            csv << "'" << s << "\t'" << st.total_levels << "\t'" << st.no_visited << "\t'" << st.kernel_ms << "\t"
                << std::endl;
            total += st.kernel_ms;
        }
        if (!sources.empty()) total /= double(sources.size());
    } else if (c.kernel == "MBFS") {
        std::cout << "MBFS kernels launching..." << std::endl;
        check(blest_query_create(dg, 128, &q));
        std::vector<int32_t> levels(size_t(n) * 128);
        blest_mbfs_opts mo{0, 128};
        for (size_t at = 0; at < sources.size(); at += 128) {
            const uint32_t k = uint32_t(std::min<size_t>(128, sources.size() - at));
            blest_bfs_stats st;
            check(blest_mbfs(q, sources.data() + at, k, levels.data(), &mo, &st));
            std::cout << "Batch of " << k << " sources - levels " << st.total_levels << " - Time took: "
                      << st.kernel_ms << " ms." << std::endl;
            total += st.kernel_ms;
        }
    } else {
        std::cout << "CC kernels launching..." << std::endl;
        check(blest_query_create(dg, 1, &q));
        std::vector<uint32_t> labels(n);
        blest_cc_stats st;
        check(blest_cc(q, 0, labels.data(), &st));
        std::cout << "Number of components identified: " << st.no_components << std::endl;
        total = st.kernel_ms;
    }
    std::cout << "Time: " << total << " ms." << std::endl;
    blest_query_free(q);
    blest_device_graph_free(dg);
    blest_bvss_free(b);
    blest_csc_free(g);
    (void)synthetic;
    return rc;
}

}  // namespace

int main(int argc, char** argv) {
    try {
        const Config c = parse(argc, argv);
        return run(c);
    } catch (const std::exception& e) {
        std::cerr << "Error: " << e.what() << std::endl;
        usage(argv[0]);
        return 1;
    }
}
