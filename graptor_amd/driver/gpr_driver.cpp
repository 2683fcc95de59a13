// gpr_driver.cpp -- command-line harness of the traversal engine, in the
// shape of Graptor's bench/driver.C:187-278: read a graph, build the
// benchmark object, time `rounds` calls of run() (and only run(),
// driver.C:245-255), then post-process and validate each round and report
// the average.  One binary serves the four benchmarks through -app.
//
//   gpr_driver -app bfslvl|bfsv|cc|pr|pr_vary|apr|kc [options] <inFile>
//
// Options follow the reference's names (driver.C:189-196, BFSLVLv.C:67-70,
// CCv.C:175-178, PRv.C:73-77, PR_vary.C:197-201, APR_vary.C:318-326,
// KC_bucket.C:60-72):
//   -b                 inFile is a b2 binary graph (the only file format)
//   -s                 the graph is symmetric
//   -rounds R          timed rounds (3)
//   -start V           BFS source vertex, original ID (0)
//   -pr:maxiter K      PR iteration cap (100)
//   -pr:outfile F      write the ranks, one per line, original order
//   -cc:outfile F      write the recoded component labels, one per line
//   -kc:outfile F      write the coreness, one per line
//   -kc:buckets B      accepted; the peeling has no open-bucket window
//   -itimes            per-iteration lines on stderr (BFSLVLv.C:90-106)
// and, without a counterpart in the reference:
//   -rmat S            generate RMAT-S on the device instead of reading a file
//   -ef E, -seed X     RMAT edge factor (16) and seed (0x47524150)
//   -reorder           relabel by decreasing degree (GraphVEBOGraptor's role)
//   -dir D             auto | push | pull | policy  (auto = m/20 threshold)
//   -threshold T       frac_threshold(T) percent (default: the benchmark's own
//                      rule, default_threshold m/20 for BFS, frac_threshold(5) for CC)
//   -fusion F          fused sparse steps: default|off|1|2 (the benchmarks'
//                      FUSION setting, BFSLVLv.C:28-40, CCv.C:47-54) or small
//                      (small levels only, no reference counterpart)
//   -deferred 0|1      the benchmarks' DEFERRED_UPDATE record forms (0: the
//                      reduction; 1: method / reduction_or_method with a
//                      maintained previous copy, BFSLVLv.C:10-12, CCv.C:26-28)
//   -policy F          network weights (JSON with a "params" array) for -dir policy
//   -pr:stages L       PR_vary / APR_vary stages fmt:tol,... with fmt
//                      f32|bf16|f16|cfp16|bf16t|scfp16 (default cfp16:1e-3,f32:1e-7
//                      for pr_vary, PR_vary.C:41-43 VARIANT 4620; f32:1e-1,scfp16:1e-7
//                      for apr, APR_vary.C:37-39 VARIANT 2046)
//   -apr:eps2 E        APR_vary's activity threshold (1e-2, APR_vary.C:913)
//   -device N          HIP device (0)
//   -tuned F           runtime tunables file (default: tuned.json beside the
//                      binary, when present; GPR_NO_TUNED=1 skips it)
//
// "Running : t" is the device time of run() in seconds, measured on the
// graph's stream; the copy of the result to the host is outside it, as the
// reference keeps its result in host memory.
#include "gpr.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <limits>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include <unistd.h>

namespace {

struct Args {
    std::map<std::string, std::string> opt;
    std::vector<std::string> pos;
    bool has(const char *k) const { return opt.count(k) != 0; }
    std::string str(const char *k, const char *d) const {
        auto it = opt.find(k);
        return it == opt.end() ? d : it->second;
    }
    long long num(const char *k, long long d) const {
        auto it = opt.find(k);
        if (it == opt.end()) return d;
        char *end = nullptr;
        long long v = std::strtoll(it->second.c_str(), &end, 0);
        if (!end || *end) throw std::runtime_error(std::string("bad value for ") + k + ": " + it->second);
        return v;
    }
    double real(const char *k, double d) const {
        auto it = opt.find(k);
        return it == opt.end() ? d : std::strtod(it->second.c_str(), nullptr);
    }
};

// Flags that take no value; every other -option consumes the next word.
const char *kFlags[] = {"-b", "-s", "-itimes", "-reorder", "-debug", "-h", "--help"};

Args parse(int argc, char **argv) {
    Args a;
    for (int i = 1; i < argc; ++i) {
        std::string w = argv[i];
        if (w.size() > 1 && w[0] == '-') {
            bool flag = std::find_if(std::begin(kFlags), std::end(kFlags),
                                     [&](const char *f) { return w == f; }) != std::end(kFlags);
            if (flag) {
                a.opt[w] = "1";
            } else {
                if (i + 1 >= argc) throw std::runtime_error("option " + w + " needs a value");
                a.opt[w] = argv[++i];
            }
        } else {
            a.pos.push_back(w);
        }
    }
    return a;
}

void usage() {
    std::cerr << "usage: gpr_driver -app bfslvl|bfsv|cc|pr|pr_vary|apr|kc [-b] [-s] [-rounds R] [-start V]\n"
                 "                  [-pr:maxiter K] [-pr:outfile F] [-cc:outfile F] [-kc:outfile F] [-itimes]\n"
                 "                  [-apr:eps2 E] [-kc:buckets B]\n"
                 "                  [-rmat S [-ef E] [-seed X]] [-reorder] [-dir auto|push|pull|policy]\n"
                 "                  [-threshold T] [-fusion default|off|1|2|small] [-deferred 0|1] [-policy F] [-pr:stages fmt:tol,...] [-device N] [-tuned F]\n"
                 "                  <inFile>\n";
}

[[noreturn]] void fail(const char *what, int rc) {
    char msg[512] = {0};
    gpr_last_error(msg, sizeof msg);
    std::cerr << "gpr_driver: " << what << " failed (" << rc << "): " << msg << "\n";
    std::exit(1);
}

#define CALL(expr)                          \
    do {                                    \
        int rc_ = (expr);                   \
        if (rc_ != GPR_OK) fail(#expr, rc_); \
    } while (0)

int contrib_fmt(const std::string &s) {
    if (s == "f32") return GPR_CONTRIB_F32;
    if (s == "bf16") return GPR_CONTRIB_BF16;
    if (s == "f16") return GPR_CONTRIB_F16;
    if (s == "cfp16") return GPR_CONTRIB_CFP16;
    if (s == "bf16t") return GPR_CONTRIB_BF16T;
    if (s == "scfp16") return GPR_CONTRIB_SCFP16;
    throw std::runtime_error("unknown contribution format " + s);
}

std::vector<gpr_pr_stage> parse_stages(const std::string &s) {
    std::vector<gpr_pr_stage> out;
    std::stringstream ss(s);
    std::string item;
    while (std::getline(ss, item, ',')) {
        size_t c = item.find(':');
        if (c == std::string::npos) throw std::runtime_error("stage '" + item + "' is not fmt:tol");
        out.push_back({contrib_fmt(item.substr(0, c)), std::strtod(item.c_str() + c + 1, nullptr)});
    }
    if (out.empty()) throw std::runtime_error("no PR stages");
    return out;
}

std::vector<float> read_policy(const std::string &path) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error("cannot read " + path);
    std::string text((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    size_t k = text.find("\"params\"");
    if (k == std::string::npos || (k = text.find('[', k)) == std::string::npos)
        throw std::runtime_error(path + " has no \"params\" array");
    std::vector<float> p;
    const char *c = text.c_str() + k + 1;
    for (;;) {
        while (*c == ' ' || *c == ',' || *c == '\n' || *c == '\r' || *c == '\t') ++c;
        if (*c == ']' || !*c) break;
        char *end = nullptr;
        p.push_back(std::strtof(c, &end));
        if (end == c) throw std::runtime_error(path + ": bad number in params");
        c = end;
    }
    return p;
}

// The "tunables" object of a tuned.json file: "name": number pairs.
void apply_tuned(const std::string &path, bool required) {
    std::ifstream f(path);
    if (!f) {
        if (required) throw std::runtime_error("cannot read " + path);
        return;
    }
    std::string text((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    size_t k = text.find("\"tunables\"");
    if (k == std::string::npos || (k = text.find('{', k)) == std::string::npos)
        throw std::runtime_error(path + " has no \"tunables\" object");
    const size_t end = text.find('}', k);
    while (true) {
        size_t q0 = text.find('"', k);
        if (q0 == std::string::npos || q0 > end) break;
        size_t q1 = text.find('"', q0 + 1);
        size_t colon = text.find(':', q1);
        if (q1 == std::string::npos || colon == std::string::npos || colon > end)
            throw std::runtime_error(path + ": malformed tunables");
        const std::string name = text.substr(q0 + 1, q1 - q0 - 1);
        char *num_end = nullptr;
        const double v = std::strtod(text.c_str() + colon + 1, &num_end);
        CALL(gpr_tune_set(name.c_str(), v));
        k = (size_t)(num_end - text.c_str());
    }
}

std::string exe_dir() {
    char buf[4096];
    ssize_t len = readlink("/proc/self/exe", buf, sizeof buf - 1);
    if (len <= 0) return ".";
    buf[len] = 0;
    std::string s(buf);
    size_t slash = s.rfind('/');
    return slash == std::string::npos ? "." : s.substr(0, slash);
}

const char *input_kind(const gpr_iter_stats &s, bool all) {
    if (all) return "true";
    return s.direction == GPR_DIR_PUSH ? "sparse" : "bit";
}

void dump_itimes(gpr_graph *g, uint64_t m, uint32_t n, bool all_true) {
    std::vector<gpr_iter_stats> tr(4096);
    int k = 0;
    CALL(gpr_trace_get(g, tr.data(), (int)tr.size(), &k));
    double active = 0.0;
    for (int i = 0; i < k && i < (int)tr.size(); ++i) {
        const gpr_iter_stats &s = tr[i];
        active += s.nactv_in;
        std::cerr << "Iteration " << i << ": " << s.ms * 1e-3 << ' ' << input_kind(s, all_true)
                  << " density: " << (double(s.nactv_in) + double(s.nacte_in)) / double(m)
                  << " nacte: " << s.nacte_in << " nactv: " << s.nactv_in
                  << " active: " << float(active) / float(n) << "\n";
    }
}

// The CSR of the graph in original IDs, for the validators.
struct HostGraph {
    uint32_t n = 0;
    uint64_t m = 0;
    std::vector<uint64_t> index;
    std::vector<uint32_t> nbrs, order;
    explicit HostGraph(gpr_graph *g) {
        int sym = 0;
        CALL(gpr_graph_info(g, &n, &m, &sym));
        index.resize((size_t)n + 1);
        nbrs.resize(m);
        order.resize(n);
        CALL(gpr_graph_download(g, index.data(), nbrs.data()));
        CALL(gpr_graph_order(g, order.data()));
    }
    template <class F> uint64_t count_edges(F &&bad) const {
        uint64_t wrong = 0;
#pragma omp parallel for schedule(dynamic, 4096) reduction(+ : wrong)
        for (int64_t u = 0; u < (int64_t)n; ++u) {
            bool w = false;
            for (uint64_t e = index[u]; e < index[u + 1] && !w; ++e) w = bad(order[u], order[nbrs[e]]);
            wrong += w;
        }
        return wrong;
    }
};

// BFSLVLv.C:265-322
void validate_bfs(gpr_graph *g, uint32_t start, const std::vector<uint32_t> &level, bool symmetric) {
    const uint32_t n = (uint32_t)level.size();
    uint32_t longest = 0, longest_v = start;
    uint64_t active = 0;
    for (uint32_t v = 0; v < n; ++v) {
        if (level[v] <= n) ++active;
        if (longest < level[v] && level[v] < n) {
            longest = level[v];
            longest_v = v;
        }
    }
    uint32_t start_int = 0, longest_int = 0;
    CALL(gpr_graph_remap_id(g, start, &start_int));
    CALL(gpr_graph_remap_id(g, longest_v, &longest_int));
    std::cout << "Longest path from " << start_int << " (original: " << start << ") : " << longest
              << " at: " << longest_int << " (" << longest_v << ")\n";
    HostGraph h(g);
    // a vertex whose neighbour is more than one level away (symmetric), or
    // one that could have been reached earlier through an in-neighbour
    uint64_t wrong = symmetric
        ? h.count_edges([&](uint32_t s, uint32_t d) {
              return (uint64_t)level[s] > (uint64_t)level[d] + 1 || (uint64_t)level[s] + 1 < level[d];
          })
        : h.count_edges([&](uint32_t s, uint32_t d) { return (uint64_t)level[s] + 1 < level[d]; });
    std::cout << "Number of vertices with nghbs not activated in time: " << wrong << (wrong == 0 ? " PASS" : " FAIL")
              << "\n";
    std::cout << "Number of activated vertices: " << active << "\n";
    std::cout << "Number of vertices: " << n << "\n";
}

// BFSv.C:205-230: no parent chain longer than the traversal's depth, every
// vertex activated at most once.  Chain lengths are found with memoised
// walks (linear time); a walk longer than n (a cycle) counts as too long.
void validate_bfsv(gpr_graph *g, uint32_t start, int iters, const std::vector<uint32_t> &parent) {
    const uint32_t n = (uint32_t)parent.size();
    const uint32_t longest = iters > 0 ? (uint32_t)(iters - 1) : 0;
    uint32_t start_int = 0;
    CALL(gpr_graph_remap_id(g, start, &start_int));
    std::cout << "Longest path from " << start_int << " (original: " << start << ") : " << longest << "\n";
    // depth[v]: parent steps from v to its root (0 for the root and for
    // unset vertices, as the reference's loop counts them); BAD for a chain
    // that leaves the vertex range or runs longer than n (a cycle)
    const uint32_t UNK = 0xffffffffu, BAD = 0xfffffffeu;
    std::vector<uint32_t> depth(n, UNK), path;
    uint64_t longer = 0, active = 0;
    for (uint32_t v = 0; v < n; ++v) {
        if (parent[v] != 0xffffffffu) ++active;
        path.clear();
        uint32_t p = v, d = 0;
        for (;;) {
            if (depth[p] != UNK) { d = depth[p]; break; }
            const uint32_t q = parent[p];
            if (q == 0xffffffffu || q == p) { depth[p] = d = 0; break; }
            if (q >= n || path.size() >= n) { d = BAD; break; }
            path.push_back(p);
            p = q;
        }
        for (size_t i = path.size(); i-- > 0;) {
            d = d == BAD ? BAD : d + 1;
            depth[path[i]] = d;
        }
        if (depth[v] == BAD || depth[v] > longest) ++longer;
    }
    std::cout << "Number of vertices with longer chains: " << longer << (longer == 0 ? " PASS" : " FAIL") << "\n";
    std::cout << "Number of activated vertices: " << active << "\n";
    std::cout << "Number of vertices: " << n << "\n";
    std::cout << "Every vertex activated at most once: " << (n >= active ? "PASS" : "FAIL") << "\n";
}

// CCv.C:604-660 with count_unique (bench/unique.h:44-128)
void validate_cc(gpr_graph *g, const std::vector<uint32_t> &ids) {
    const uint32_t n = (uint32_t)ids.size();
    std::vector<uint32_t> count(n, 0);
    uint64_t others = 0;
    for (uint32_t v = 0; v < n; ++v) {
        if (ids[v] >= n) ++others;
        else ++count[ids[v]];
    }
    uint32_t nunique = 0, largest = 0, smallest = n, most = 0, least = 0;
    bool first = true;
    uint64_t sum = others;
    for (uint32_t j = 0; j < n; ++j) {
        if (!count[j]) continue;
        ++nunique;
        sum += count[j];
        smallest = std::min(smallest, j);
        largest = std::max(largest, j);
        if (first || count[j] > count[most]) most = j;
        if (first || count[j] < count[least]) least = j;
        first = false;
    }
    std::cout << "Unique values: " << nunique << "\n"
              << "Check: All vertices accounted for? " << (sum == n ? "PASS" : "FAIL") << " (s=" << sum << '/' << n
              << '/' << n << ")\n"
              << "Largest value: " << largest << "\n"
              << "Smallest value: " << smallest << "\n"
              << "Least frequent value: " << least << " count: " << (nunique ? count[least] : 0) << "\n"
              << "Most frequent value: " << most << " count: " << (nunique ? count[most] : 0) << "\n";
    if (nunique && largest - smallest + 1 < nunique)
        std::cout << "FAIL: largest-smallest span (" << largest - smallest << " less then number of unique values ("
                  << nunique << ")\n";
    if (others) std::cout << "No values out of range? FAIL (others=" << others << ")\n";
    HostGraph h(g);
    uint64_t wrong = h.count_edges([&](uint32_t s, uint32_t d) { return ids[s] != ids[d]; });
    if (wrong == 0)
        std::cerr << "Neighbour check: PASS\n";
    else
        std::cerr << "Neighbour check: FAIL; #vertices: " << wrong << "\n";
}

// KC_bucket.C:332-460: no coreness above the degree, coreness 0 exactly for
// isolated vertices, at least coreness-many neighbours of at least the same
// coreness, at most coreness-many of a higher one.
void validate_kc(gpr_graph *g, const std::vector<uint32_t> &core) {
    HostGraph h(g);
    const uint32_t n = h.n;
    std::vector<uint32_t> deg(n);
    for (uint32_t u = 0; u < n; ++u) deg[h.order[u]] = (uint32_t)(h.index[u + 1] - h.index[u]);
    uint64_t above = 0, zero = 0;
    for (uint32_t v = 0; v < n; ++v) {
        above += core[v] > deg[v];
        zero += (core[v] == 0) != (deg[v] == 0);
    }
    uint64_t few = 0, low = 0;
#pragma omp parallel for schedule(dynamic, 4096) reduction(+ : few, low)
    for (int64_t u = 0; u < (int64_t)n; ++u) {
        const uint32_t v = h.order[u], cv = core[v];
        uint64_t ge = 0, gt = 0;
        for (uint64_t e = h.index[u]; e < h.index[u + 1]; ++e) {
            const uint32_t cu = core[h.order[h.nbrs[e]]];
            ge += cu >= cv;
            gt += cu > cv;
        }
        few += ge < cv;
        low += gt > cv;
    }
    if (above || zero || few || low)
        std::cerr << "Validation failed: " << above << " above degree, " << zero << " zero/degree mismatches, "
                  << few << " with too few neighbours, " << low << " with too low k-core number\n";
    else
        std::cerr << "Validation successful\n";
}

template <class T> void write_lines(const std::string &path, const std::vector<T> &v, int precision) {
    std::ofstream f(path);
    if (!f) throw std::runtime_error("cannot write " + path);
    f.precision(precision);
    for (const T &x : v) f << x << '\n';
}

int run(const Args &a) {
    const std::string app = a.str("-app", "");
    if (app != "bfslvl" && app != "bfsv" && app != "cc" && app != "pr" && app != "pr_vary" && app != "apr" &&
        app != "kc") {
        usage();
        return 2;
    }
    const bool symmetric = a.has("-s");
    const long rounds = (long)a.num("-rounds", 3);
    if (rounds < 1) throw std::runtime_error("-rounds must be at least 1");

    gpr_run_opts o = {};
    const std::string dir = a.str("-dir", "auto");
    o.direction = dir == "auto" ? GPR_DIR_AUTO
                : dir == "push" ? GPR_DIR_PUSH
                : dir == "pull" ? GPR_DIR_PULL
                : dir == "policy" ? GPR_DIR_POLICY
                : -1;
    if (o.direction < 0) throw std::runtime_error("unknown -dir " + dir);
    o.threshold_pct = (float)a.real("-threshold", 0.0);   // 0: the benchmark's own rule
    o.max_iter = (int)a.num("-pr:maxiter", 100);
    o.tol = 1e-7;
    o.trace = a.has("-itimes") ? 1 : 0;
    const std::string fu = a.str("-fusion", "default");
    o.fusion = fu == "default" ? 0 : fu == "off" ? -1 : fu == "1" ? 1 : fu == "2" ? 2 : fu == "small" ? 3 : -2;
    if (o.fusion == -2) throw std::runtime_error("unknown -fusion " + fu);
    o.deferred_update = (int)a.num("-deferred", 0) != 0;
    if (a.has("-pr:infile")) throw std::runtime_error("-pr:infile is not supported: runs start from x = 1/n");

    CALL(gpr_init((int)a.num("-device", 0)));
    if (a.has("-tuned")) apply_tuned(a.str("-tuned", ""), true);
    else if (!std::getenv("GPR_NO_TUNED")) apply_tuned(exe_dir() + "/tuned.json", false);
    gpr_graph *g = nullptr;
    if (a.has("-rmat")) {
        CALL(gpr_graph_generate_rmat((uint32_t)a.num("-rmat", 20), (uint32_t)a.num("-ef", 16), 0.57, 0.19, 0.19,
                                     (uint64_t)a.num("-seed", 0x47524150LL), &g));
    } else {
        if (a.pos.empty()) {
            usage();
            return 2;
        }
        if (!a.has("-b")) throw std::runtime_error("driver supports only binary mode files (-b)");
        CALL(gpr_graph_load_b2(a.pos.back().c_str(), symmetric ? 1 : 0, &g));
    }
    if (a.has("-reorder")) CALL(gpr_graph_reorder(g, GPR_ORDER_DEGREE));
    if (a.has("-policy")) {
        std::vector<float> p = read_policy(a.str("-policy", ""));
        CALL(gpr_policy_set(g, p.data(), (int)p.size()));
    }
    uint32_t n = 0;
    uint64_t m = 0;
    int sym = 0;
    CALL(gpr_graph_info(g, &n, &m, &sym));
    std::cerr << "graph: n=" << n << " m=" << m << (sym ? " symmetric" : " directed") << "\n";

    const uint32_t start = (uint32_t)a.num("-start", 0);
    std::vector<gpr_pr_stage> stages =
        app == "pr" ? std::vector<gpr_pr_stage>{{GPR_CONTRIB_F32, 1e-7}}
                    : parse_stages(a.str("-pr:stages", app == "apr" ? "f32:1e-1,scfp16:1e-7" : "cfp16:1e-3,f32:1e-7"));
    const double eps2 = a.real("-apr:eps2", 1e-2);
    const bool pr_like = app == "pr" || app == "pr_vary" || app == "apr";
    std::vector<uint32_t> out_u32(app == "bfslvl" || app == "bfsv" || app == "cc" || app == "kc" ? n : 0);
    std::vector<float> out_f32(app == "pr" || app == "pr_vary" ? n : 0);
    std::vector<double> out_f64(app == "apr" ? n : 0);
    double total = 0.0;
    uint32_t kmax = 0;
    for (long r = 0; r < rounds; ++r) {
        gpr_run_info info = {};
        if (app == "bfslvl") CALL(gpr_bfslvl(g, start, &o, out_u32.data(), &info));
        else if (app == "bfsv") CALL(gpr_bfsv(g, start, &o, out_u32.data(), &info));
        else if (app == "cc") CALL(gpr_ccv(g, &o, out_u32.data(), &info));
        else if (app == "kc") CALL(gpr_kcore(g, &o, out_u32.data(), &kmax, &info));
        else if (app == "apr")
            CALL(gpr_apr_vary(g, stages.data(), (int)stages.size(), eps2, &o, out_f64.data(), &info));
        else CALL(gpr_pr_vary(g, stages.data(), (int)stages.size(), &o, out_f32.data(), &info));
        const double sec = info.ms * 1e-3;
        total += sec;
        std::cout << "Running : " << std::setprecision(6) << sec << std::endl;
        std::cout << "Iterations : " << info.iters << " push: " << info.push_iters << " edges: " << info.edges
                  << " GTEPS: " << (sec > 0 ? double(info.edges) / sec * 1e-9 : 0.0) << "\n";
        // post_process
        if (o.trace && app != "kc") dump_itimes(g, m, n, app == "pr" || app == "pr_vary");
        if (pr_like && info.iters >= o.max_iter) std::cerr << "WARNING: took all allowed iterations\n";
        if (app == "cc" && a.has("-cc:outfile")) write_lines(a.str("-cc:outfile", ""), out_u32, 10);
        if (app == "kc") {
            std::cout << "Largest core: " << kmax << "\n";
            if (a.has("-kc:outfile")) write_lines(a.str("-kc:outfile", ""), out_u32, 10);
        }
        if ((app == "pr" || app == "pr_vary") && a.has("-pr:outfile"))
            write_lines(a.str("-pr:outfile", ""), out_f32, std::numeric_limits<float>::max_digits10);
        if (app == "apr" && a.has("-pr:outfile"))
            write_lines(a.str("-pr:outfile", ""), out_f64, std::numeric_limits<double>::max_digits10);
        // validate
        if (app == "bfslvl") validate_bfs(g, start, out_u32, sym != 0);
        else if (app == "bfsv") validate_bfsv(g, start, info.iters, out_u32);
        else if (app == "cc") validate_cc(g, out_u32);
        else if (app == "kc") validate_kc(g, out_u32);
        else if (app == "apr") {
            for (uint32_t i = 0; i < 10 && i < n; ++i) std::cout << "PR[" << i << "]: " << out_f64[i] << "\n";
            std::cout << "delta: " << info.delta << " sum: " << info.sum << "\n";
        } else {
            for (uint32_t i = 0; i < 10 && i < n; ++i) std::cout << "PR[" << i << "]: " << out_f32[i] << "\n";
            std::cout << "delta: " << info.delta << " sum: " << info.sum << "\n";
        }
    }
    std::cout << "Average : " << std::setprecision(6) << total / rounds << std::endl;
    CALL(gpr_graph_destroy(g));
    return 0;
}

} // namespace

int main(int argc, char **argv) {
    try {
        Args a = parse(argc, argv);
        if (a.has("-h") || a.has("--help")) {
            usage();
            return 0;
        }
        return run(a);
    } catch (const std::exception &e) {
        std::cerr << "gpr_driver: " << e.what() << "\n";
        return 2;
    }
}
